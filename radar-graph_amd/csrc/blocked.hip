// Blocked PageRank engine: the all-vertices push of PageRank.C:122
// (p_next[d] += p_curr[s]/outdeg(s) over every arc) reorganised for MI355X
// memory, with no global atomics.
//
// On this part every global atomic is performed at the memory side (the L2
// forwards it), so a push that scatters FP64 atomics over an n-vertex
// accumulator is bound by memory-side atomic throughput, and a gather over
// in-edges by 128-byte line fetches for 8-byte reads.  The blocked layout
// keeps both sides on chip:
//   * destinations are cut into blocks of 2^kVB vertices whose FP64
//     accumulators fit in one workgroup's LDS; a workgroup ("unit") owns an
//     edge range of one block and accumulates with LDS atomics;
//   * a block with more than kUnitEdges in-edges (the hub blocks of a
//     degree-sorted graph) is split over several units, each with its own
//     private LDS copy, merged by the vertex phase -- the reference's hub
//     duplication (hubdup PageRank.C:45-73) at workgroup granularity;
//   * within a block, edges are ordered by source segment (2^kSB sources),
//     so the contrib[s] reads of the units running at the same time fall in
//     few segments and hit in each XCD's L2, which keeps its own copy.
// Per arc the kernel streams a 4-byte source id and a 2-byte local
// destination; per unit it writes its 2^kVB partial sums.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "blocked.h"
#include "device.h"
#include "queries.h"

namespace rg {

constexpr int kMaxVB = 14;                       // at most 16K destinations per block (128 KB LDS)
constexpr int kUnitThreads = 1024;
constexpr int kMaxXcds = 16;                     // per-XCD work queues (k_blk_pr_qx)

void BlockedCSC::release() {
  if (src) (void)hipFree(src);
  if (vl) (void)hipFree(vl);
  if (unit_lo) (void)hipFree(unit_lo);
  if (block_unit0) (void)hipFree(block_unit0);
  if (unit_block) (void)hipFree(unit_block);
  if (multi) (void)hipFree(multi);
  if (mchunk) (void)hipFree(mchunk);
  if (by_src) (void)hipFree(by_src);
  if (xcd_shares) (void)hipFree(xcd_shares);
  for (uint32_t* q : {pk, seg_rel, seg_hi, unit_seg0})
    if (q) (void)hipFree(q);
  pk = seg_rel = seg_hi = unit_seg0 = nullptr;
  nsegs = 0;
  sbits = 0;
  pack_asked = false;
  src = nullptr; vl = nullptr; unit_lo = nullptr; block_unit0 = nullptr; unit_block = nullptr;
  multi = nullptr;
  mchunk = nullptr;
  nmchunk = 0;
  by_src = nullptr;
  xcd_shares = nullptr;
  nsplit = 0;
  nxcd = 0;
  built_for = nullptr;
  nmulti = 0;
}

void blocked_free(void* p) {
  BlockedCSC* b = static_cast<BlockedCSC*>(p);
  if (b) {
    b->release();
    delete b;
  }
}

namespace {

// One wavefront per destination: key = block << bs | segment << vb | local,
// with bs = vb + (32 - sb) so the segment field (32 - sb bits) never
// overlaps the block field; the key needs 64 - sb bits.
__global__ void k_blk_keys(const uint64_t* in_off, const uint32_t* in_nbr, int64_t n, int vb, int sb,
                           int bs, uint64_t* keys, uint32_t* vals) {
  const int lane = (int)__lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  const int64_t lmask = (int64_t(1) << vb) - 1;
  for (int64_t v = wid; v < n; v += nw) {
    const uint64_t b = in_off[v], e = in_off[v + 1];
    const uint64_t hi = ((uint64_t)(v >> vb) << bs) | (uint64_t)(v & lmask);
    for (uint64_t j = b + lane; j < e; j += kWave) {
      uint32_t s = in_nbr[j];
      keys[j] = hi | ((uint64_t)(s >> sb) << vb);
      vals[j] = s;
    }
  }
}

// First source of every unit (kNone for an empty unit).
__global__ void k_unit_first_src(const uint32_t* src, const uint64_t* unit_lo, int64_t nunits, uint32_t* out) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nunits;
       u += (int64_t)gridDim.x * blockDim.x)
    out[u] = unit_lo[u] < unit_lo[u + 1] ? src[unit_lo[u]] : kNone;
}

__global__ void k_blk_split(const uint64_t* keys, int64_t m, int vb, int bs, uint16_t* vl, uint64_t* block_lo) {
  const uint64_t lmask = (uint64_t(1) << vb) - 1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    vl[i] = (uint16_t)(k & lmask);
    uint64_t blk = k >> bs;
    if (i == 0 || (keys[i - 1] >> bs) != blk) block_lo[blk] = (uint64_t)i;
  }
}

// One workgroup per unit: LDS accumulators for the unit's block, streamed
// (source, local destination) pairs, partial sums written out at the end.
// The two streams are read once (nontemporal loads, so they do not evict
// the L2-resident source segment); four edges per lane are in flight.
template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// Sums contrib[src] into acc[local destination] over the unit's edges
// [lo, hi) (LDS atomics; four edges per lane in flight with UNROLL = 4).
template <int THREADS, bool NT, int UNROLL>
__device__ __forceinline__ void blk_accumulate(const uint32_t* __restrict__ src, const uint16_t* __restrict__ vl,
                                               uint64_t lo, uint64_t hi, const double* __restrict__ contrib,
                                               double* acc) {
  uint64_t e = lo + threadIdx.x;
  if (UNROLL == 4) {
    for (; e + 3 * THREADS < hi; e += 4 * THREADS) {
      const uint32_t s0 = ld_stream<NT>(src + e);
      const uint32_t s1 = ld_stream<NT>(src + e + THREADS);
      const uint32_t s2 = ld_stream<NT>(src + e + 2 * THREADS);
      const uint32_t s3 = ld_stream<NT>(src + e + 3 * THREADS);
      const uint32_t l0 = ld_stream<NT>(vl + e);
      const uint32_t l1 = ld_stream<NT>(vl + e + THREADS);
      const uint32_t l2 = ld_stream<NT>(vl + e + 2 * THREADS);
      const uint32_t l3 = ld_stream<NT>(vl + e + 3 * THREADS);
      const double c0 = contrib[s0], c1 = contrib[s1], c2 = contrib[s2], c3 = contrib[s3];
      atomicAdd(&acc[l0], c0);
      atomicAdd(&acc[l1], c1);
      atomicAdd(&acc[l2], c2);
      atomicAdd(&acc[l3], c3);
    }
  }
  for (; e < hi; e += THREADS) {
    const uint32_t s = ld_stream<NT>(src + e);
    const uint32_t l = ld_stream<NT>(vl + e);
    atomicAdd(&acc[l], contrib[s]);
  }
}

// The packed arcs and their source segments (BlockedCSC::pk, seg_rel,
// seg_hi, unit_seg0, sbits); pk == nullptr: the layout keeps src + vl.
struct PackedArcs {
  const uint32_t* pk;
  const uint32_t* seg_rel;
  const uint32_t* seg_hi;
  const uint32_t* unit_seg0;
  int sbits;
};

// A unit's source segments staged in LDS: s_rel[j] = first arc of segment j
// (relative to the unit's first arc), s_hi[j] its high source bits;
// s_rel[count] = UINT32_MAX.  Synchronised by the caller's next barrier.
template <int THREADS>
__device__ __forceinline__ void stage_segments(const PackedArcs& pa, int64_t u, uint32_t* s_rel, uint32_t* s_hi) {
  const uint32_t g0 = pa.unit_seg0[u], ns = pa.unit_seg0[u + 1] - g0;
  for (uint32_t i = threadIdx.x; i < ns; i += THREADS) {
    s_rel[i] = pa.seg_rel[g0 + i];
    s_hi[i] = pa.seg_hi[g0 + i];
  }
  if (threadIdx.x == 0) s_rel[ns] = 0xffffffffu;
}

// blk_accumulate over packed arcs: arc r (relative) holds source
// s_hi[j] << sbits | pk >> VB for the segment j holding r.  A thread's arcs
// move forward (slot after slot, step after step), so one segment cursor
// serves them all; it keeps the next segment's first arc and the current
// high bits in registers and reads the staged table only when an arc
// crosses into a later segment (RMAT-26 PageRank, 20 iterations: 105.5 ms
// against 106.8 with both read from LDS per arc).  UNROLL (1 or 4) arcs per
// lane in flight (8 measured slower: 106.4 ms).
template <int VB, int THREADS, bool NT, int UNROLL>
__device__ __forceinline__ void blk_accumulate_pk(const uint32_t* __restrict__ pk, uint64_t lo, uint64_t hi,
                                                  const uint32_t* s_rel, const uint32_t* s_hi, int sbits,
                                                  const double* __restrict__ contrib, double* acc) {
  constexpr uint32_t LM = (1u << VB) - 1;
  const uint32_t len = (uint32_t)(hi - lo);
  const uint32_t* p = pk + lo;
  uint32_t r = threadIdx.x;
  int c = 0;
  uint32_t nx = s_rel[1], hb = s_hi[0];
  if (UNROLL > 1) {
    for (; r + (UNROLL - 1) * THREADS < len; r += UNROLL * THREADS) {
      uint32_t w[UNROLL], sh[UNROLL];
#pragma unroll
      for (int k = 0; k < UNROLL; ++k) w[k] = ld_stream<NT>(p + r + k * THREADS);
#pragma unroll
      for (int k = 0; k < UNROLL; ++k) {
        const uint32_t rk = r + k * THREADS;
        while (rk >= nx) {
          ++c;
          nx = s_rel[c + 1];
          hb = s_hi[c];
        }
        sh[k] = hb;
      }
      double x[UNROLL];
#pragma unroll
      for (int k = 0; k < UNROLL; ++k) x[k] = contrib[(sh[k] << sbits) | (w[k] >> VB)];
#pragma unroll
      for (int k = 0; k < UNROLL; ++k) atomicAdd(&acc[w[k] & LM], x[k]);
    }
  }
  for (; r < len; r += THREADS) {
    const uint32_t w = ld_stream<NT>(p + r);
    while (r >= nx) {
      ++c;
      nx = s_rel[c + 1];
      hb = s_hi[c];
    }
    atomicAdd(&acc[w & LM], contrib[(hb << sbits) | (w >> VB)]);
  }
}

// PR_Vertex_F (PageRank.C:55) for vertex v from its in-edge sum a, fused
// with the next iteration's contribution: returns |old - new| (the L1 term).
__device__ __forceinline__ double pr_finalize(int64_t v, double a, double* rank, const uint64_t* off,
                                              double* contrib_next, double damping, double added) {
  const double nw = __dadd_rn(__dmul_rn(damping, a), added);
  const double d = fabs(rank[v] - nw);
  rank[v] = nw;
  const uint64_t deg = off[v + 1] - off[v];
  contrib_next[v] = deg ? nw / (double)(uint32_t)deg : 0.0;
  return d;
}

// Sum of x over the workgroup (THREADS a multiple of 64), valid in thread 0.
template <int THREADS>
__device__ __forceinline__ double wg_sum(double x) {
  __shared__ double red[THREADS / kWave];
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
  if (__lane_id() == 0) red[threadIdx.x / kWave] = x;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < THREADS / kWave; ++i) t += red[i];
  return t;
}

// One unit of a PageRank iteration on the blocked layout: the unit sums its
// in-edges into the LDS block acc; a unit that holds its block alone then
// applies the vertex step to the block's vertices itself (new rank in place,
// L1 term, next contribution into the other contribution buffer); units of
// split blocks write partials, which k_blk_pr_merge adds in unit order.
// diff_part[u] = the unit's L1 terms (0 for split-block units).  Every
// thread of the workgroup calls it; acc is free again when it returns.
template <int VB, int THREADS, bool NT, int UNROLL>
__device__ __forceinline__ void pr_unit(int64_t u, double* acc, const uint32_t* __restrict__ src,
                                        const uint16_t* __restrict__ vl, const uint64_t* __restrict__ unit_lo,
                                        const uint32_t* __restrict__ unit_block,
                                        const uint32_t* __restrict__ block_unit0, const double* __restrict__ contrib,
                                        double* __restrict__ contrib_next, double* rank,
                                        const uint64_t* __restrict__ off, int64_t n, double* __restrict__ part,
                                        double* diff_part, double damping, double added, const PackedArcs& pa,
                                        uint32_t* s_rel, uint32_t* s_hi) {
  constexpr int BV = 1 << VB;
  for (int i = threadIdx.x; i < BV; i += THREADS) acc[i] = 0.0;
  if (pa.pk) stage_segments<THREADS>(pa, u, s_rel, s_hi);
  __syncthreads();
  if (pa.pk)
    blk_accumulate_pk<VB, THREADS, NT, UNROLL>(pa.pk, unit_lo[u], unit_lo[u + 1], s_rel, s_hi, pa.sbits, contrib,
                                               acc);
  else
    blk_accumulate<THREADS, NT, UNROLL>(src, vl, unit_lo[u], unit_lo[u + 1], contrib, acc);
  __syncthreads();
  const uint32_t b = unit_block[u];
  if (block_unit0[b + 1] - block_unit0[b] != 1) {
    double* out = part + u * (int64_t)BV;
    for (int i = threadIdx.x; i < BV; i += THREADS) out[i] = acc[i];
    if (threadIdx.x == 0) diff_part[u] = 0.0;
    __syncthreads();
    return;
  }
  // The vertex step, four vertices per thread at a time: their loads are
  // issued together (one LDS block per CU, so this tail is not hidden by
  // another workgroup's edges).
  constexpr int FU = 4;
  const int64_t base = (int64_t)b << VB;
  double diff = 0.0;
  for (int i0 = threadIdx.x; i0 < BV; i0 += THREADS * FU) {
    double r[FU];
    uint64_t o0[FU], o1[FU];
#pragma unroll
    for (int k = 0; k < FU; ++k) {
      const int i = i0 + k * THREADS;
      const int64_t v = base + i;
      if (i < BV && v < n) {
        r[k] = rank[v];
        o0[k] = off[v];
        o1[k] = off[v + 1];
      }
    }
#pragma unroll
    for (int k = 0; k < FU; ++k) {
      const int i = i0 + k * THREADS;
      const int64_t v = base + i;
      if (i < BV && v < n) {
        const double nw = __dadd_rn(__dmul_rn(damping, 0.0 + acc[i]), added);
        diff += fabs(r[k] - nw);
        rank[v] = nw;
        const uint64_t deg = o1[k] - o0[k];
        contrib_next[v] = deg ? nw / (double)(uint32_t)deg : 0.0;
      }
    }
  }
  diff = wg_sum<THREADS>(diff);
  if (threadIdx.x == 0) diff_part[u] = diff;
  __syncthreads();
}

// One PageRank iteration, one workgroup per unit (order: workgroup -> unit,
// nullptr = the layout's order).
template <int VB, int THREADS, bool NT, int UNROLL>
__global__ __launch_bounds__(THREADS) void k_blk_pr_it(const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ src,
                                                       const uint16_t* __restrict__ vl,
                                                       const uint64_t* __restrict__ unit_lo,
                                                       const uint32_t* __restrict__ unit_block,
                                                       const uint32_t* __restrict__ block_unit0,
                                                       const double* __restrict__ contrib,
                                                       double* __restrict__ contrib_next, double* rank,
                                                       const uint64_t* __restrict__ off, int64_t n,
                                                       double* __restrict__ part, double* diff_part,
                                                       const int64_t* done, double damping, double added,
                                                       const int64_t* shares, int nxcd, unsigned long long* heads,
    PackedArcs pa) {
  __shared__ double acc[1 << VB];
  __shared__ uint32_t s_rel[kMaxSegs + 1], s_hi[kMaxSegs];
  if (*done) return;
  const int64_t u = order ? order[blockIdx.x] : blockIdx.x;
  pr_unit<VB, THREADS, NT, UNROLL>(u, acc, src, vl, unit_lo, unit_block, block_unit0, contrib, contrib_next, rank,
                                   off, n, part, diff_part, damping, added, pa, s_rel, s_hi);
}

// XCD of the calling wave (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id() { return (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15); }

// Takes the next unit for the calling workgroup (thread 0 only) from the
// per-XCD shares of `order` (nullptr: the layout's order): share x =
// order[shares[2x], shares[2x + 1]) for x < X, share X the common tail; heads[x] counts what share x has handed
// out.  A workgroup drains its own XCD's share, then the other XCDs' in
// turn, then the common share; *cur is the share it takes from (starts at
// its XCD), *tried the XCD shares found empty.  Returns -1 when all are
// empty.
__device__ __forceinline__ int64_t take_unit(const uint32_t* __restrict__ order, const int64_t* __restrict__ shares,
                                             unsigned long long* heads, int X, int* cur, int* tried) {
  for (;;) {
    const int c = *cur;
    const int64_t q = shares[2 * c] + (int64_t)atomicAdd(&heads[c], 1ull);
    if (q < shares[2 * c + 1]) return order ? (int64_t)order[q] : q;
    if (c == X) return -1;
    *tried += 1;
    *cur = *tried < X ? (c + 1) % X : X;
  }
}

// One PageRank iteration on persistent workgroups (as many as fit on the
// chip at once) taking units from a work queue, so no CU waits for a
// workgroup launch and no XCD idles while another still holds units (with
// one workgroup per unit, XCDs took the same number of units but finished
// up to 800 us apart in a 6.1 ms iteration).  By default the queue hands
// out the split blocks' units by first source, so units of different hub
// blocks reading the same source range run at the same time, then the
// single-unit blocks.  With per-XCD shares, that sequence is cut into X
// contiguous shares, one per XCD, and a workgroup whose XCD's share is
// empty takes from the others'.  k_pr_finish_b clears the queue heads.
template <int VB, int THREADS, bool NT, int UNROLL>
__global__ __launch_bounds__(THREADS) void k_blk_pr_qx(const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ src,
                                                       const uint16_t* __restrict__ vl,
                                                       const uint64_t* __restrict__ unit_lo,
                                                       const uint32_t* __restrict__ unit_block,
                                                       const uint32_t* __restrict__ block_unit0,
                                                       const double* __restrict__ contrib,
                                                       double* __restrict__ contrib_next, double* rank,
                                                       const uint64_t* __restrict__ off, int64_t n,
                                                       double* __restrict__ part, double* diff_part,
                                                       const int64_t* done, double damping, double added,
                                                       const int64_t* shares, int nxcd, unsigned long long* heads,
    PackedArcs pa) {
  __shared__ double acc[1 << VB];
  __shared__ uint32_t s_rel[kMaxSegs + 1], s_hi[kMaxSegs];
  __shared__ long long s_u;
  if (*done) return;
  int cur = xcc_id() % nxcd, tried = 0;
  for (;;) {
    if (threadIdx.x == 0) s_u = take_unit(order, shares, heads, nxcd, &cur, &tried);
    __syncthreads();
    const int64_t u = (int64_t)s_u;
    if (u < 0) return;
    pr_unit<VB, THREADS, NT, UNROLL>(u, acc, src, vl, unit_lo, unit_block, block_unit0, contrib, contrib_next,
                                     rank, off, n, part, diff_part, damping, added, pa, s_rel, s_hi);
  }
}

// Sum of vertex l's partials over units [u, u1) of a split block: unit u + k
// goes into accumulator k mod K, so that K loads are in flight (a hub block
// holds up to ~10^3 units, and the merge waits for its longest chain of
// dependent round trips), and the K sums are added pairwise in a fixed
// order.
template <int VB, int K>
__device__ __forceinline__ double sum_units(const double* __restrict__ part, uint32_t u, uint32_t u1, int l) {
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  for (; u + K <= u1; u += K) {
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] += part[((int64_t)(u + k) << VB) + l];
  }
  for (int k = 0; u < u1; ++u, ++k) a[k] += part[((int64_t)u << VB) + l];
#pragma unroll
  for (int w = K / 2; w >= 1; w /= 2)
#pragma unroll
    for (int k = 0; k < w; ++k) a[k] += a[k + w];
  return a[0];
}

// Units' partials summed per vertex in flight at a time by the merges.
constexpr int kMergeLoads = 16;

// Split blocks: a vertex's sum is its units' partials (sum_units), then the
// vertex step.  Grid: nmulti * (2^VB / 256) workgroups of 256;
// diff_part[workgroup] = its L1 terms.
template <int VB>
__global__ __launch_bounds__(256) void k_blk_pr_merge(const uint32_t* __restrict__ multi,
                                                      const uint32_t* __restrict__ block_unit0,
                                                      const double* __restrict__ part, double* rank,
                                                      const uint64_t* __restrict__ off,
                                                      double* __restrict__ contrib_next, int64_t n,
                                                      double* diff_part, const int64_t* done, double damping,
                                                      double added) {
  constexpr int PER = (1 << VB) / 256;
  if (*done) return;
  const uint32_t b = multi[blockIdx.x / PER];
  const int l = (int)(blockIdx.x % PER) * 256 + threadIdx.x;
  const int64_t v = ((int64_t)b << VB) + l;
  double diff = 0.0;
  if (v < n) {
    const double sum = sum_units<VB, kMergeLoads>(part, block_unit0[b], block_unit0[b + 1], l);
    diff = pr_finalize(v, sum, rank, off, contrib_next, damping, added);
  }
  diff = wg_sum<256>(diff);
  if (threadIdx.x == 0) diff_part[blockIdx.x] = diff;
}

// vertexFilter of one PageRankDelta round for vertex v, from its in-edge sum
// ns, its old p and out-degree (PR_Vertex_F_FirstRound, PageRankDelta.C:51-64,
// in the first round; PR_Vertex_F, 66-78, after): updates p[v], writes the
// next round's quotient Delta/outdeg (0 when v is not active) into
// contrib_next[v] and returns whether v is active; |Delta| goes to l1.
// Multiplies and adds are rounded separately, as the reference's host code
// computes them.
__device__ __forceinline__ bool prd_finalize(int64_t v, double ns, double pv, uint64_t deg, double* p,
                                             double* contrib_next, const PrdStep& st, double& l1) {
  double dl;
  bool act;
  if (st.first) {
    dl = __dadd_rn(__dmul_rn(st.damping, ns), st.added);
    pv = __dadd_rn(pv, dl);
    dl = __dsub_rn(dl, st.one_over_n);
    act = fabs(dl) > __dmul_rn(st.eps2, pv);
    p[v] = pv;
  } else {
    dl = __dmul_rn(ns, st.damping);
    act = fabs(dl) > __dmul_rn(st.eps2, pv);
    if (act) p[v] = __dadd_rn(pv, dl);
  }
  l1 += fabs(dl);
  contrib_next[v] = (act && deg) ? dl / (double)(uint32_t)deg : 0.0;
  return act;
}

// Appends the vertices whose act[k] is set (FU per thread) to out_ids, in
// thread order within the workgroup: a wave scan of the per-thread counts,
// the waves' totals added up by thread 0 and one atomicAdd on *count per
// call.  Every thread of the workgroup calls it.
template <int THREADS, int FU>
__device__ __forceinline__ void wg_append(const bool* act, const int64_t* vs, uint32_t* __restrict__ out_ids,
                                          int64_t* count, uint32_t* s_cnt, unsigned long long* s_base) {
  int c = 0;
#pragma unroll
  for (int k = 0; k < FU; ++k) c += act[k] ? 1 : 0;
  const int lane = __lane_id();
  int x = c;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  const int w = threadIdx.x / kWave;
  if (lane == kWave - 1) s_cnt[w] = (uint32_t)x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < THREADS / kWave; ++i) {
      const uint32_t y = s_cnt[i];
      s_cnt[i] = t;
      t += y;
    }
    *s_base = t ? atomicAdd(reinterpret_cast<unsigned long long*>(count), (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  uint64_t pos = *s_base + s_cnt[w] + (uint32_t)(x - c);
#pragma unroll
  for (int k = 0; k < FU; ++k)
    if (act[k]) out_ids[pos++] = (uint32_t)vs[k];
  __syncthreads();   // s_cnt and s_base are reused by the next call
}

// Sums l1 and deg over the workgroup into *l1_out and *deg_out (thread 0).
template <int THREADS>
__device__ __forceinline__ void wg_finish(double l1, uint64_t deg, double* l1_out, uint64_t* deg_out) {
  __shared__ double r1[THREADS / kWave];
  __shared__ uint64_t r2[THREADS / kWave];
  for (int o = kWave / 2; o > 0; o >>= 1) {
    l1 += __shfl_xor(l1, o);
    deg += __shfl_xor(deg, o);
  }
  if (__lane_id() == 0) {
    r1[threadIdx.x / kWave] = l1;
    r2[threadIdx.x / kWave] = deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    uint64_t d = 0;
    for (int i = 0; i < THREADS / kWave; ++i) {
      a += r1[i];
      d += r2[i];
    }
    *l1_out = a;
    *deg_out = d;
  }
}

// One unit of a dense PageRankDelta round on the blocked layout, edge phase
// and vertexFilter fused as in pr_unit: a unit that holds its block alone
// applies the vertex step to the block (next quotients into contrib_next,
// active vertices appended to out_ids, their count added to *count); units
// of split blocks write partials for k_blk_prd_merge.  l1_part[u] and
// deg_part[u] = the unit's |Delta| sum and its active vertices' out-degree
// sum (0 for split-block units).  Every thread of the workgroup calls it;
// acc, s_cnt and s_base are free again when it returns.
template <int VB, int THREADS, bool NT, int UNROLL>
__device__ __forceinline__ void prd_unit(int64_t u, double* acc, uint32_t* s_cnt, unsigned long long* s_base,
                                         const uint32_t* __restrict__ src, const uint16_t* __restrict__ vl,
                                         const uint64_t* __restrict__ unit_lo,
                                         const uint32_t* __restrict__ unit_block,
                                         const uint32_t* __restrict__ block_unit0,
                                         const double* __restrict__ contrib, double* __restrict__ contrib_next,
                                         double* p, const uint64_t* __restrict__ off, int64_t n,
                                         double* __restrict__ part, double* l1_part, uint64_t* deg_part,
                                         uint32_t* __restrict__ out_ids, int64_t* count, const PrdStep& st,
                                         const PackedArcs& pa, uint32_t* s_rel, uint32_t* s_hi) {
  constexpr int BV = 1 << VB;
  constexpr int FU = 4;
  for (int i = threadIdx.x; i < BV; i += THREADS) acc[i] = 0.0;
  if (pa.pk) stage_segments<THREADS>(pa, u, s_rel, s_hi);
  __syncthreads();
  if (pa.pk)
    blk_accumulate_pk<VB, THREADS, NT, UNROLL>(pa.pk, unit_lo[u], unit_lo[u + 1], s_rel, s_hi, pa.sbits, contrib,
                                               acc);
  else
    blk_accumulate<THREADS, NT, UNROLL>(src, vl, unit_lo[u], unit_lo[u + 1], contrib, acc);
  __syncthreads();
  const uint32_t b = unit_block[u];
  if (block_unit0[b + 1] - block_unit0[b] != 1) {
    double* out = part + u * (int64_t)BV;
    for (int i = threadIdx.x; i < BV; i += THREADS) out[i] = acc[i];
    if (threadIdx.x == 0) {
      l1_part[u] = 0.0;
      deg_part[u] = 0;
    }
    __syncthreads();
    return;
  }
  const int64_t base = (int64_t)b << VB;
  double l1 = 0.0;
  uint64_t dsum = 0;
  // The same number of passes for every thread (wg_append synchronizes).
  for (int pass = 0; pass < (BV + THREADS * FU - 1) / (THREADS * FU); ++pass) {
    const int i0 = pass * THREADS * FU + threadIdx.x;
    double pv[FU];
    uint64_t o0[FU], o1[FU];
    int64_t vs[FU];
    bool act[FU];
#pragma unroll
    for (int k = 0; k < FU; ++k) {
      const int i = i0 + k * THREADS;
      vs[k] = base + i;
      act[k] = false;
      if (i < BV && vs[k] < n) {
        pv[k] = p[vs[k]];
        o0[k] = off[vs[k]];
        o1[k] = off[vs[k] + 1];
      }
    }
#pragma unroll
    for (int k = 0; k < FU; ++k) {
      const int i = i0 + k * THREADS;
      if (i < BV && vs[k] < n) {
        const uint64_t deg = o1[k] - o0[k];
        act[k] = prd_finalize(vs[k], 0.0 + acc[i], pv[k], deg, p, contrib_next, st, l1);
        if (act[k]) dsum += deg;
      }
    }
    wg_append<THREADS, FU>(act, vs, out_ids, count, s_cnt, s_base);
  }
  wg_finish<THREADS>(l1, dsum, &l1_part[u], &deg_part[u]);
  __syncthreads();
}

// One dense PageRankDelta round, one workgroup per unit (order: workgroup ->
// unit, nullptr = the layout's order).
template <int VB, int THREADS, bool NT, int UNROLL>
__global__ __launch_bounds__(THREADS) void k_blk_prd_it(const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ src,
                                                        const uint16_t* __restrict__ vl,
                                                        const uint64_t* __restrict__ unit_lo,
                                                        const uint32_t* __restrict__ unit_block,
                                                        const uint32_t* __restrict__ block_unit0,
                                                        const double* __restrict__ contrib,
                                                        double* __restrict__ contrib_next, double* p,
                                                        const uint64_t* __restrict__ off, int64_t n,
                                                        double* __restrict__ part, double* l1_part,
                                                        uint64_t* deg_part, uint32_t* __restrict__ out_ids,
                                                        int64_t* count, const int64_t* done, PrdStep st,
                                                        const int64_t* shares, int nxcd, unsigned long long* heads,
    PackedArcs pa) {
  __shared__ double acc[1 << VB];
  __shared__ uint32_t s_rel[kMaxSegs + 1], s_hi[kMaxSegs];
  __shared__ uint32_t s_cnt[THREADS / kWave];
  __shared__ unsigned long long s_base;
  if (*done) return;
  const int64_t u = order ? order[blockIdx.x] : blockIdx.x;
  prd_unit<VB, THREADS, NT, UNROLL>(u, acc, s_cnt, &s_base, src, vl, unit_lo, unit_block, block_unit0, contrib,
                                    contrib_next, p, off, n, part, l1_part, deg_part, out_ids, count, st, pa, s_rel,
                                    s_hi);
}

// One dense PageRankDelta round on persistent workgroups taking units from
// per-XCD shares (take_unit, as k_blk_pr_qx); k_prd_round_sum clears the
// queue heads.
template <int VB, int THREADS, bool NT, int UNROLL>
__global__ __launch_bounds__(THREADS) void k_blk_prd_qx(const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ src,
                                                        const uint16_t* __restrict__ vl,
                                                        const uint64_t* __restrict__ unit_lo,
                                                        const uint32_t* __restrict__ unit_block,
                                                        const uint32_t* __restrict__ block_unit0,
                                                        const double* __restrict__ contrib,
                                                        double* __restrict__ contrib_next, double* p,
                                                        const uint64_t* __restrict__ off, int64_t n,
                                                        double* __restrict__ part, double* l1_part,
                                                        uint64_t* deg_part, uint32_t* __restrict__ out_ids,
                                                        int64_t* count, const int64_t* done, PrdStep st,
                                                        const int64_t* shares, int nxcd, unsigned long long* heads,
    PackedArcs pa) {
  __shared__ double acc[1 << VB];
  __shared__ uint32_t s_rel[kMaxSegs + 1], s_hi[kMaxSegs];
  __shared__ uint32_t s_cnt[THREADS / kWave];
  __shared__ unsigned long long s_base;
  __shared__ long long s_u;
  if (*done) return;
  int cur = xcc_id() % nxcd, tried = 0;
  for (;;) {
    if (threadIdx.x == 0) s_u = take_unit(order, shares, heads, nxcd, &cur, &tried);
    __syncthreads();
    const int64_t u = (int64_t)s_u;
    if (u < 0) return;
    prd_unit<VB, THREADS, NT, UNROLL>(u, acc, s_cnt, &s_base, src, vl, unit_lo, unit_block, block_unit0, contrib,
                                      contrib_next, p, off, n, part, l1_part, deg_part, out_ids, count, st, pa,
                                      s_rel, s_hi);
  }
}

// Split blocks of a dense PageRankDelta round: a vertex's sum is its units'
// partials (sum_units, as in k_blk_pr_merge), then the vertex step and the
// append of k_blk_prd_it, four vertices per thread (one append, so one
// atomic, and one pair of round sums per 1024 vertices: k_prd_round_sum
// reads every workgroup's pair).  Grid: nmulti * (2^VB / 1024) workgroups
// of 256; l1_part / deg_part[workgroup] = its sums.
template <int VB>
__global__ __launch_bounds__(256) void k_blk_prd_merge(const uint32_t* __restrict__ multi,
                                                       const uint32_t* __restrict__ block_unit0,
                                                       const double* __restrict__ part, double* p,
                                                       const uint64_t* __restrict__ off,
                                                       double* __restrict__ contrib_next, int64_t n,
                                                       double* l1_part, uint64_t* deg_part,
                                                       uint32_t* __restrict__ out_ids, int64_t* count,
                                                       const int64_t* done, PrdStep st) {
  constexpr int FU = 4;
  constexpr int PER = (1 << VB) / (256 * FU);
  __shared__ uint32_t s_cnt[256 / kWave];
  __shared__ unsigned long long s_base;
  if (*done) return;
  const uint32_t b = multi[blockIdx.x / PER];
  double l1 = 0.0;
  uint64_t dsum = 0;
  bool act[FU];
  int64_t vs[FU];
#pragma unroll
  for (int k = 0; k < FU; ++k) {
    const int l = (int)(blockIdx.x % PER) * (256 * FU) + k * 256 + threadIdx.x;
    const int64_t v = ((int64_t)b << VB) + l;
    vs[k] = v;
    act[k] = false;
    if (v < n) {
      const double sum = sum_units<VB, kMergeLoads>(part, block_unit0[b], block_unit0[b + 1], l);
      const uint64_t deg = off[v + 1] - off[v];
      act[k] = prd_finalize(v, sum, p[v], deg, p, contrib_next, st, l1);
      if (act[k]) dsum += deg;
    }
  }
  wg_append<256, FU>(act, vs, out_ids, count, s_cnt, &s_base);
  wg_finish<256>(l1, dsum, &l1_part[blockIdx.x], &deg_part[blockIdx.x]);
}

// The round's L1 norm and active out-degree sum from the workgroups' parts
// (one workgroup, a fixed order).
__global__ __launch_bounds__(256) void k_prd_round_sum(const double* l1_part, const uint64_t* deg_part,
                                                       int64_t nparts, double* l1_out, int64_t* deg_sum) {
  double a = 0.0;
  uint64_t d = 0;
  for (int64_t i = threadIdx.x; i < nparts; i += 256) {
    a += l1_part[i];
    d += deg_part[i];
  }
  wg_finish<256>(a, d, l1_out, reinterpret_cast<uint64_t*>(deg_sum));
}

int grid_n(rg_ctx* ctx, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 16));
}

}  // namespace

namespace {

// Packed-arc build (BlockedCSC::pk): segment starts flagged (every unit's
// first arc, every arc whose source segment differs from the previous
// arc's), numbered by a scan, then the tables and the 4-byte arcs.
__global__ void k_flag_units(const uint64_t* unit_lo, int64_t nunits, uint64_t m, uint32_t* flags) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nunits; u += (int64_t)gridDim.x * blockDim.x)
    if (unit_lo[u] < unit_lo[u + 1]) flags[unit_lo[u]] = 1;
}

__global__ void k_flag_segments(const uint32_t* src, int64_t m, int sbits, uint32_t* flags) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + 1; e < m; e += (int64_t)gridDim.x * blockDim.x)
    if ((src[e] >> sbits) != (src[e - 1] >> sbits)) flags[e] = 1;
}

__global__ void k_fill_segments(const uint32_t* flags, const uint32_t* incl, const uint32_t* src, int64_t m,
                                const uint64_t* unit_lo, int64_t nunits, int sbits, uint32_t* seg_rel,
                                uint32_t* seg_hi) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
    if (!flags[e]) continue;
    // The unit holding arc e: the last u with unit_lo[u] <= e.
    int64_t a = 0, b = nunits - 1;
    while (a < b) {
      const int64_t mid = (a + b + 1) / 2;
      if (unit_lo[mid] <= (uint64_t)e) a = mid;
      else b = mid - 1;
    }
    const uint32_t j = incl[e] - 1;
    seg_rel[j] = (uint32_t)((uint64_t)e - unit_lo[a]);
    seg_hi[j] = src[e] >> sbits;
  }
}

__global__ void k_unit_segments(const uint64_t* unit_lo, int64_t nunits, const uint32_t* flags, const uint32_t* incl,
                                uint64_t m, uint32_t nsegs, uint32_t* unit_seg0) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u <= nunits; u += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = unit_lo[u];
    unit_seg0[u] = lo < m ? incl[lo] - flags[lo] : nsegs;
  }
}

__global__ void k_pack_arcs(const uint32_t* src, const uint16_t* vl, int64_t m, int sbits, int vb, uint32_t* pk) {
  const uint32_t mask = sbits ? (sbits >= 32 ? 0xffffffffu : (1u << sbits) - 1) : 0u;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    pk[e] = ((src[e] & mask) << vb) | (uint32_t)vl[e];
}

// Builds the packed arcs of B when the graph allows (2^sbits-source
// segments, at most kMaxSegs per unit, and sbits + vb <= 32 bits); src and
// vl are freed then.  Otherwise B keeps src + vl.
int pack_arcs(rg_ctx* ctx, BlockedCSC* B, int64_t n, int64_t m, int vb) {
  int lg = 0;
  while ((int64_t(1) << lg) < n) ++lg;
  const int sbits = std::max(0, lg - 10);   // (n - 1) >> sbits < kMaxSegs
  if (m <= 0 || sbits + vb > 32) return RG_OK;
  hipStream_t s = ctx->stream;
  DevBuf flags, incl, tmp;
  RG_TRY(flags.ensure(sizeof(uint32_t) * m));
  RG_TRY(incl.ensure(sizeof(uint32_t) * m));
  RG_HIP(hipMemsetAsync(flags.p, 0, sizeof(uint32_t) * m, s));
  hipLaunchKernelGGL(k_flag_units, grid_n(ctx, B->nunits), 256, 0, s, B->unit_lo, B->nunits, (uint64_t)m,
                     flags.as<uint32_t>());
  hipLaunchKernelGGL(k_flag_segments, grid_n(ctx, m), 256, 0, s, B->src, m, sbits, flags.as<uint32_t>());
  size_t tb = 0;
  RG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, flags.as<uint32_t>(), incl.as<uint32_t>(), (int)m, s));
  RG_TRY(tmp.ensure(tb));
  RG_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, flags.as<uint32_t>(), incl.as<uint32_t>(), (int)m, s));
  uint32_t nsegs = 0;
  RG_TRY(ctx_copy(ctx, &nsegs, incl.as<uint32_t>() + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
  RG_HIP(hipMalloc(&B->seg_rel, sizeof(uint32_t) * std::max<uint32_t>(nsegs, 1)));
  RG_HIP(hipMalloc(&B->seg_hi, sizeof(uint32_t) * std::max<uint32_t>(nsegs, 1)));
  RG_HIP(hipMalloc(&B->unit_seg0, sizeof(uint32_t) * (B->nunits + 1)));
  RG_HIP(hipMalloc(&B->pk, sizeof(uint32_t) * m));
  hipLaunchKernelGGL(k_fill_segments, grid_n(ctx, m), 256, 0, s, flags.as<uint32_t>(), incl.as<uint32_t>(), B->src,
                     m, B->unit_lo, B->nunits, sbits, B->seg_rel, B->seg_hi);
  hipLaunchKernelGGL(k_unit_segments, grid_n(ctx, B->nunits + 1), 256, 0, s, B->unit_lo, B->nunits,
                     flags.as<uint32_t>(), incl.as<uint32_t>(), (uint64_t)m, nsegs, B->unit_seg0);
  hipLaunchKernelGGL(k_pack_arcs, grid_n(ctx, m), 256, 0, s, B->src, B->vl, m, sbits, vb, B->pk);
  RG_HIP(hipGetLastError());
  // Every unit within kMaxSegs segments (holds by the choice of sbits).
  std::vector<uint32_t> us(B->nunits + 1);
  RG_TRY(ctx_copy(ctx, us.data(), B->unit_seg0, sizeof(uint32_t) * us.size(), hipMemcpyDeviceToHost));
  for (int64_t u = 0; u < B->nunits; ++u)
    if (us[u + 1] - us[u] > (uint32_t)kMaxSegs)
      return set_error(RG_EDEVICE, "blocked layout: a unit spans more than kMaxSegs source segments");
  (void)hipFree(B->src);
  (void)hipFree(B->vl);
  B->src = nullptr;
  B->vl = nullptr;
  B->nsegs = nsegs;
  B->sbits = sbits;
  return RG_OK;
}

}  // namespace

int blocked_build(rg_graph* g, int vb, int sb, void** slot, BlockedCSC** out, bool pack) {
  BlockedCSC* B = static_cast<BlockedCSC*>(*slot);
  if (B && B->built_for == g->in_off && B->vb == vb && B->sb == sb && B->pack_asked == pack) {
    *out = B;
    return RG_OK;
  }
  if (g->m > (int64_t)INT32_MAX) return set_error(RG_EUNSUPPORTED, "blocked layout: more than 2^31 arcs");
  rg_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  if (borrowed(g, B)) B = nullptr;   // the base's layout stays as it is
  if (!B) {
    B = new BlockedCSC;
    *slot = B;
  }
  B->release();
  const int64_t n = g->n, m = g->m;
  const int64_t nblocks = (n + (int64_t(1) << vb) - 1) >> vb;
  DevBuf keys, skeys, vals, tmp, blo;
  RG_TRY(keys.ensure(sizeof(uint64_t) * std::max<int64_t>(m, 1)));
  RG_TRY(skeys.ensure(sizeof(uint64_t) * std::max<int64_t>(m, 1)));
  RG_TRY(vals.ensure(sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  RG_HIP(hipMalloc(&B->src, sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  RG_HIP(hipMalloc(&B->vl, sizeof(uint16_t) * std::max<int64_t>(m, 1)));
  RG_TRY(blo.ensure(sizeof(uint64_t) * (nblocks + 1)));
  RG_HIP(hipMemsetAsync(blo.p, 0xff, sizeof(uint64_t) * (nblocks + 1), s));
  const int bs = vb + (32 - sb);
  hipLaunchKernelGGL(k_blk_keys, grid_n(ctx, n), 256, 0, s, g->in_off, g->in_nbr, n, vb, sb, bs,
                     keys.as<uint64_t>(), vals.as<uint32_t>());
  const int end_bit = bs + std::max(1, (int)(64 - __builtin_clzll((unsigned long long)nblocks)));
  size_t tb = 0;
  RG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint64_t>(), skeys.as<uint64_t>(),
                                            vals.as<uint32_t>(), B->src, (int)m, 0, end_bit, s));
  RG_TRY(tmp.ensure(tb));
  RG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<uint64_t>(), skeys.as<uint64_t>(),
                                            vals.as<uint32_t>(), B->src, (int)m, 0, end_bit, s));
  hipLaunchKernelGGL(k_blk_split, grid_n(ctx, m), 256, 0, s, skeys.as<uint64_t>(), m, vb, bs, B->vl,
                     blo.as<uint64_t>());
  RG_HIP(hipGetLastError());
  std::vector<uint64_t> block_lo(nblocks + 1);
  RG_HIP(hipMemcpyAsync(block_lo.data(), blo.p, sizeof(uint64_t) * (nblocks + 1), hipMemcpyDeviceToHost, s));
  RG_HIP(hipStreamSynchronize(s));
  keys.release();
  skeys.release();
  vals.release();
  tmp.release();
  // Blocks without edges start where the next non-empty block starts.
  block_lo[nblocks] = (uint64_t)m;
  for (int64_t b = nblocks - 1; b >= 0; --b)
    if (block_lo[b] == ~0ull) block_lo[b] = block_lo[b + 1];
  std::vector<uint64_t> ulo;
  std::vector<uint32_t> bu0(nblocks + 1);
  for (int64_t b = 0; b < nblocks; ++b) {
    bu0[b] = (uint32_t)ulo.size();
    uint64_t lo = block_lo[b], hi = block_lo[b + 1];
    if (hi == lo) {
      ulo.push_back(lo);   // an empty unit keeps every block's partials defined
      continue;
    }
    for (uint64_t e = lo; e < hi; e += (uint64_t)kUnitEdges) ulo.push_back(e);
  }
  bu0[nblocks] = (uint32_t)ulo.size();
  B->nunits = (int64_t)ulo.size();
  ulo.push_back((uint64_t)m);
  B->nblocks = nblocks;
  std::vector<uint32_t> ub(B->nunits), multi, mchunk;
  for (int64_t b = 0; b < nblocks; ++b) {
    for (uint32_t u = bu0[b]; u < bu0[b + 1]; ++u) ub[u] = (uint32_t)b;
    if (bu0[b + 1] - bu0[b] > 1) {
      multi.push_back((uint32_t)b);
      for (uint32_t u = bu0[b]; u < bu0[b + 1]; u += kMergeChunk) mchunk.push_back(u);
    }
  }
  B->nmulti = (int64_t)multi.size();
  B->nmchunk = (int64_t)mchunk.size();
  RG_HIP(hipMalloc(&B->unit_lo, sizeof(uint64_t) * ulo.size()));
  RG_HIP(hipMalloc(&B->block_unit0, sizeof(uint32_t) * bu0.size()));
  RG_HIP(hipMalloc(&B->unit_block, sizeof(uint32_t) * ub.size()));
  RG_HIP(hipMalloc(&B->multi, sizeof(uint32_t) * std::max<size_t>(1, multi.size())));
  RG_TRY(ctx_copy(ctx, B->unit_lo, ulo.data(), sizeof(uint64_t) * ulo.size(), hipMemcpyHostToDevice));
  RG_TRY(ctx_copy(ctx, B->block_unit0, bu0.data(), sizeof(uint32_t) * bu0.size(), hipMemcpyHostToDevice));
  RG_TRY(ctx_copy(ctx, B->unit_block, ub.data(), sizeof(uint32_t) * ub.size(), hipMemcpyHostToDevice));
  RG_HIP(hipMalloc(&B->mchunk, sizeof(uint32_t) * std::max<size_t>(1, mchunk.size())));
  if (!multi.empty()) {
    RG_TRY(ctx_copy(ctx, B->multi, multi.data(), sizeof(uint32_t) * multi.size(), hipMemcpyHostToDevice));
    RG_TRY(ctx_copy(ctx, B->mchunk, mchunk.data(), sizeof(uint32_t) * mchunk.size(), hipMemcpyHostToDevice));
  }
  // The order by first source and its per-XCD shares (BlockedCSC::by_src).
  {
    DevBuf fs;
    RG_TRY(fs.ensure(sizeof(uint32_t) * B->nunits));
    hipLaunchKernelGGL(k_unit_first_src, grid_n(ctx, B->nunits), 256, 0, s, B->src, B->unit_lo, B->nunits,
                       fs.as<uint32_t>());
    RG_HIP(hipGetLastError());
    std::vector<uint32_t> first(B->nunits);
    RG_TRY(ctx_copy(ctx, first.data(), fs.p, sizeof(uint32_t) * B->nunits, hipMemcpyDeviceToHost));
    std::vector<uint32_t> split, single;
    for (int64_t b = 0; b < nblocks; ++b) {
      if (bu0[b + 1] - bu0[b] > 1)
        for (uint32_t u = bu0[b]; u < bu0[b + 1]; ++u) split.push_back(u);
      else
        single.push_back(bu0[b]);
    }
    std::stable_sort(split.begin(), split.end(), [&](uint32_t a, uint32_t c) { return first[a] < first[c]; });
    const int64_t X = std::min(kMaxXcds, std::max(1, ctx->num_xcds)), P = (int64_t)split.size();
    const int64_t q = P / X, r = P % X;
    std::vector<uint32_t> seq(split);
    seq.insert(seq.end(), single.begin(), single.end());
    std::vector<int64_t> shares(2 * (X + 1));
    for (int64_t x = 0; x < X; ++x) {
      shares[2 * x] = x * q + std::min(x, r);
      shares[2 * x + 1] = (x + 1) * q + std::min(x + 1, r);
    }
    shares[2 * X] = P;
    shares[2 * X + 1] = B->nunits;
    // Then the one-queue form: every unit in one share, an empty common one.
    shares.insert(shares.end(), {0, B->nunits, B->nunits, B->nunits});
    RG_HIP(hipMalloc(&B->by_src, sizeof(uint32_t) * std::max<int64_t>(1, B->nunits)));
    RG_HIP(hipMalloc(&B->xcd_shares, sizeof(int64_t) * shares.size()));
    RG_TRY(ctx_copy(ctx, B->by_src, seq.data(), sizeof(uint32_t) * B->nunits, hipMemcpyHostToDevice));
    RG_TRY(ctx_copy(ctx, B->xcd_shares, shares.data(), sizeof(int64_t) * shares.size(), hipMemcpyHostToDevice));
    B->nsplit = P;
    B->nxcd = (int)X;
  }
  if (pack) RG_TRY(pack_arcs(ctx, B, n, m, vb));
  B->pack_asked = pack;
  B->built_for = g->in_off;
  B->vb = vb;
  B->sb = sb;
  *out = B;
  return RG_OK;
}

// ---------------------------------------------------------------------------
// PageRank on the blocked layout.  The vertex phase sums the partials of
// every unit of v's block in unit order (deterministic), then applies
// PR_Vertex_F / the L1 term / the reset as the push path does.

namespace {

__global__ void k_pr_init_b(const uint64_t* off, int64_t n, double* rank, double* contrib, double one_over_n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t deg = off[v + 1] - off[v];
    rank[v] = one_over_n;
    contrib[v] = deg ? one_over_n / (double)(uint32_t)deg : 0.0;
  }
}

// L1 norm of the iteration from the per-workgroup terms (one workgroup of
// 1024 threads), its history entry, the convergence flag, and the work
// queue heads of k_blk_pr_qx cleared for the next iteration.  The
// iteration number is counted on the device (*iter), so a captured loop
// holds no per-iteration argument.
__global__ __launch_bounds__(1024) void k_pr_finish_b(const double* partial, int nparts, double* l1_hist,
                                                      int64_t* iter, int64_t* done, double epsilon,
                                                      unsigned long long* queue, int nqueue) {
  if ((int)threadIdx.x < nqueue) queue[threadIdx.x] = 0;   // the next iteration's work queues start empty
  if (*done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
  s = wg_sum<1024>(s);
  if (threadIdx.x == 0) {
    const int64_t it = *iter + 1;
    *iter = it;
    l1_hist[it - 1] = s;
    if (s < epsilon) *done = it;
  }
}

}  // namespace

namespace {
// Destination-block bits (default 2^14 vertices) and source-segment bits:
// 1..24 -> 2^bits sources per segment; 0 (default) or < 0 = order by source
// id itself (measured best: units then sweep the source range in step and
// their reads share L2 lines).
void layout_bits(const rg_opts* o, int* vb, int* sb) {
  *vb = (o && o->dst_block_bits >= 10 && o->dst_block_bits <= kMaxVB) ? o->dst_block_bits : kMaxVB;
  *sb = (o && o->src_segment_bits > 0 && o->src_segment_bits <= 24) ? o->src_segment_bits : 0;
}
}  // namespace

// How a blocked edge pass runs its units (rg_opts unit_order / unit_queue):
// persistent workgroups on a work queue (default) or one workgroup per unit;
// units by first source in one queue (default), by first source cut into
// per-XCD shares, or in the layout's order.  Measured on symmetric RMAT-26
// (degree-sorted), ms per iteration: one queue by source 5.87, per-XCD
// shares 5.84-5.90, queue in layout order 5.98, one workgroup per unit in
// layout order 6.27-6.31 (per-XCD static shares without a queue: 8.1, the
// XCDs finish far apart).  The queue heads are context counters
// 40 .. 40 + nxcd, zero when a pass starts.
struct UnitSchedule {
  bool queue = true;
  const uint32_t* order = nullptr;
  const int64_t* shares = nullptr;
  int nxcd = 1;
  int grid = 0;
  unsigned long long* heads = nullptr;
};

UnitSchedule unit_schedule(rg_ctx* ctx, const BlockedCSC* B, const rg_opts* o, int vb, int threads) {
  UnitSchedule u;
  const int uo = o ? o->unit_order : 0;
  u.queue = !(o && o->unit_queue == 1);
  u.order = uo == 2 ? nullptr : B->by_src;
  u.nxcd = uo == 1 ? B->nxcd : 1;
  u.shares = uo == 1 ? B->xcd_shares : B->xcd_shares + 2 * (B->nxcd + 1);
  // The workgroups that fit on the chip at once (LDS-bound: one per CU for
  // 2^14-vertex blocks).
  const int64_t lds_fit = (int64_t)ctx->lds_per_cu / ((int64_t(8) << vb) + 256);
  const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(2048 / threads, lds_fit));
  u.grid = u.queue ? (int)std::min<int64_t>(B->nunits, (int64_t)ctx->num_cus * per_cu) : (int)B->nunits;
  u.heads = reinterpret_cast<unsigned long long*>(ctx->counts.as<int64_t>() + 40);
  return u;
}

int pagerank_prepare(rg_graph* g, const rg_opts* o) {
  int vb, sb;
  layout_bits(o, &vb, &sb);
  BlockedCSC* B = nullptr;
  return blocked_build(g, vb, sb, &g->blocked, &B, true);
}

// One edge pass of the blocked engine: per-unit partial sums of contrib[s]
// over each destination block's in-edges into unit_part (grown as needed).
// The caller adds a block's partials in unit order: *vb_out and
// *block_unit0_out describe the layout.
int blocked_prd_round(rg_graph* g, const rg_opts* o, const double* contrib, double* contrib_next, double* p,
                      const PrdStep& st, uint32_t* out_ids, int64_t* info, double* l1, const int64_t* zero,
                      DevBuf& work) {
  rg_ctx* ctx = g->ctx;
  int vb, sb;
  layout_bits(o, &vb, &sb);
  BlockedCSC* B = nullptr;
  RG_TRY(blocked_build(g, vb, sb, &g->blocked, &B, true));
  const int64_t mgrid = B->nmulti * ((int64_t(1) << vb) / 1024);
  const int64_t nparts = B->nunits + mgrid;
  const size_t unit_doubles = (size_t)B->nunits << vb;
  RG_TRY(work.ensure(sizeof(double) * (unit_doubles + 2 * nparts)));
  double* part = work.as<double>();
  double* l1_part = part + unit_doubles;
  uint64_t* deg_part = reinterpret_cast<uint64_t*>(l1_part + nparts);
  const int threads = (o && o->wg_size > 0 && o->wg_size <= 512) ? 512 : kUnitThreads;
  // Arcs in flight per lane: 0 (default) = 4 on a degree-sorted graph,
  // where the hub blocks' long units dominate (PageRankDelta on RMAT-26,
  // before its vertex step was fused: 42.7 ms against 49.1), 1 on other id orders (PageRank on directed
  // RMAT-24 original ids: 118 GTEPS against 91).
  const bool u4 = o && o->unroll ? o->unroll == 4 : g->degree_sorted;
  const UnitSchedule us = unit_schedule(ctx, B, o, vb, threads);
  void (*kern)(const uint32_t*, const uint32_t*, const uint16_t*, const uint64_t*, const uint32_t*,
               const uint32_t*, const double*, double*, double*, const uint64_t*, int64_t, double*, double*,
               uint64_t*, uint32_t*, int64_t*, const int64_t*, PrdStep, const int64_t*, int, unsigned long long*,
               PackedArcs);
  void (*merge)(const uint32_t*, const uint32_t*, const double*, double*, const uint64_t*, double*, int64_t,
                double*, uint64_t*, uint32_t*, int64_t*, const int64_t*, PrdStep);
#define RG_PICK(VB)                                                                                   \
  do {                                                                                                \
    if (us.queue)                                                                                     \
      kern = threads == 512 ? (u4 ? k_blk_prd_qx<VB, 512, true, 4> : k_blk_prd_qx<VB, 512, true, 1>)   \
                            : (u4 ? k_blk_prd_qx<VB, 1024, true, 4> : k_blk_prd_qx<VB, 1024, true, 1>); \
    else                                                                                              \
      kern = threads == 512 ? (u4 ? k_blk_prd_it<VB, 512, true, 4> : k_blk_prd_it<VB, 512, true, 1>)   \
                            : (u4 ? k_blk_prd_it<VB, 1024, true, 4> : k_blk_prd_it<VB, 1024, true, 1>); \
    merge = k_blk_prd_merge<VB>;                                                                      \
  } while (0)
  if (vb == 14) RG_PICK(14);
  else if (vb == 13) RG_PICK(13);
  else if (vb == 12) RG_PICK(12);
  else if (vb == 11) RG_PICK(11);
  else RG_PICK(10);
#undef RG_PICK
  RG_HIP(hipMemsetAsync(us.heads, 0, sizeof(int64_t) * (us.nxcd + 1), ctx->stream));
  hipEvent_t* ev = timing_begin(ctx);
  hipLaunchKernelGGL(kern, us.grid, threads, 0, ctx->stream, us.order, B->src, B->vl, B->unit_lo, B->unit_block,
                     B->block_unit0, contrib, contrib_next, p, g->off, g->n, part, l1_part, deg_part, out_ids,
                     info, zero, st, us.shares, us.nxcd, us.heads,
                     PackedArcs{B->pk, B->seg_rel, B->seg_hi, B->unit_seg0, B->sbits});
  if (mgrid)
    hipLaunchKernelGGL(merge, (int)mgrid, 256, 0, ctx->stream, B->multi, B->block_unit0, part, p, g->off,
                       contrib_next, g->n, l1_part + B->nunits, deg_part + B->nunits, out_ids, info, zero, st);
  if (ev) RG_HIP(hipEventRecord(ev[1], ctx->stream));
  hipLaunchKernelGGL(k_prd_round_sum, 1, 256, 0, ctx->stream, l1_part, deg_part, nparts, l1, info + 1);
  RG_HIP(hipGetLastError());
  return RG_OK;
}

int pagerank_blocked(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                     int64_t* iters_out, double* l1_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  int vb, sb;
  layout_bits(o, &vb, &sb);
  const int threads = (o && o->wg_size > 0 && o->wg_size <= 512) ? 512 : kUnitThreads;
  BlockedCSC* B = nullptr;
  RG_TRY(blocked_build(g, vb, sb, &g->blocked, &B, true));
  const UnitSchedule us = unit_schedule(ctx, B, o, vb, threads);
  void (*kern)(const uint32_t*, const uint32_t*, const uint16_t*, const uint64_t*, const uint32_t*,
               const uint32_t*, const double*, double*, double*, const uint64_t*, int64_t, double*, double*,
               const int64_t*, double, double, const int64_t*, int, unsigned long long*, PackedArcs);
  void (*merge)(const uint32_t*, const uint32_t*, const double*, double*, const uint64_t*, double*, int64_t,
                double*, const int64_t*, double, double);
  const bool u4 = o && o->unroll ? o->unroll == 4 : g->degree_sorted;   // default as in blocked_prd_round
#define RG_PICK(VB)                                                                                   \
  do {                                                                                                \
    if (us.queue)                                                                                     \
      kern = threads == 512 ? (u4 ? k_blk_pr_qx<VB, 512, true, 4> : k_blk_pr_qx<VB, 512, true, 1>)     \
                            : (u4 ? k_blk_pr_qx<VB, 1024, true, 4> : k_blk_pr_qx<VB, 1024, true, 1>); \
    else                                                                                              \
      kern = threads == 512 ? (u4 ? k_blk_pr_it<VB, 512, true, 4> : k_blk_pr_it<VB, 512, true, 1>)     \
                            : (u4 ? k_blk_pr_it<VB, 1024, true, 4> : k_blk_pr_it<VB, 1024, true, 1>); \
    merge = k_blk_pr_merge<VB>;                                                                       \
  } while (0)
  if (vb == 14) RG_PICK(14);
  else if (vb == 13) RG_PICK(13);
  else if (vb == 12) RG_PICK(12);
  else if (vb == 11) RG_PICK(11);
  else RG_PICK(10);
#undef RG_PICK
  DevBuf &b_rank = ctx->scratch[0], &b_contrib = ctx->scratch[1], &b_contrib2 = ctx->scratch[4],
         &b_part = ctx->scratch[3], &b_unit = ctx->scratch[2];
  RG_TRY(b_rank.ensure(sizeof(double) * n));
  RG_TRY(b_contrib.ensure(sizeof(double) * n));
  RG_TRY(b_contrib2.ensure(sizeof(double) * n));
  RG_TRY(b_unit.ensure(sizeof(double) * ((size_t)B->nunits << vb)));
  double* unit_part = b_unit.as<double>();
  // L1 terms: one per unit, then one per merge workgroup; then the history.
  const int64_t mgrid = B->nmulti * ((int64_t(1) << vb) / 256);
  const int64_t nparts = B->nunits + mgrid;
  RG_TRY(b_part.ensure(sizeof(double) * (nparts + std::max<int64_t>(maxiters, 1) + 8)));
  double* rank = b_rank.as<double>();
  double* contrib[2] = {b_contrib.as<double>(), b_contrib2.as<double>()};
  double* partial = b_part.as<double>();
  double* l1_hist = partial + nparts;
  int64_t* done = ctx->counts.as<int64_t>() + 16;
  int64_t* iter_dev = ctx->counts.as<int64_t>() + 18;
  RG_HIP(hipMemsetAsync(done, 0, sizeof(int64_t), ctx->stream));
  RG_HIP(hipMemsetAsync(iter_dev, 0, sizeof(int64_t), ctx->stream));
  const int nheads = us.nxcd + 1;
  RG_HIP(hipMemsetAsync(us.heads, 0, sizeof(int64_t) * nheads, ctx->stream));
  const double damping = 0.85, epsilon = 0.0000001;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * (1 / (double)n);
  hipLaunchKernelGGL(k_pr_init_b, grid_n(ctx, n), 256, 0, ctx->stream, g->off, n, rank, contrib[0], one_over_n);
  RG_HIP(hipGetLastError());
  hipEvent_t t0, t1;
  RG_HIP(hipEventCreate(&t0));
  RG_HIP(hipEventCreate(&t1));
  RG_HIP(hipEventRecord(t0, ctx->stream));
  // Iteration it reads contrib[(it - 1) & 1] and writes the other; once
  // converged, the later iterations' kernels return at once (*done).
  auto iterations = [&](bool timed) -> int {
    for (int64_t it = 1; it <= maxiters; ++it) {
      const double* cur = contrib[(it - 1) & 1];
      double* nxt = contrib[it & 1];
      hipEvent_t* ev = timed ? timing_begin(ctx) : nullptr;
      hipLaunchKernelGGL(kern, us.grid, threads, 0, ctx->stream, us.order, B->src, B->vl, B->unit_lo,
                         B->unit_block, B->block_unit0, cur, nxt, rank, g->off, n, unit_part, partial, done, damping,
                         added, us.shares, us.nxcd, us.heads,
                         PackedArcs{B->pk, B->seg_rel, B->seg_hi, B->unit_seg0, B->sbits});
      if (mgrid)
        hipLaunchKernelGGL(merge, (int)mgrid, 256, 0, ctx->stream, B->multi, B->block_unit0, unit_part, rank,
                           g->off, nxt, n, partial + B->nunits, done, damping, added);
      if (ev) RG_HIP(hipEventRecord(ev[1], ctx->stream));
      hipLaunchKernelGGL(k_pr_finish_b, 1, 1024, 0, ctx->stream, partial, (int)nparts, l1_hist, iter_dev, done,
                         epsilon, us.heads, nheads);
      RG_HIP(hipGetLastError());
    }
    return RG_OK;
  };
  if (ctx->graphs && !ctx->timing && maxiters > 0) {
    // The whole loop as one graph, captured once for these arguments.
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)B->src, (uint64_t)(uintptr_t)B->vl, (uint64_t)(uintptr_t)B->unit_lo,
        (uint64_t)(uintptr_t)B->unit_block, (uint64_t)(uintptr_t)B->block_unit0, (uint64_t)(uintptr_t)B->multi,
        (uint64_t)B->nunits, (uint64_t)mgrid, (uint64_t)(uintptr_t)merge,
        (uint64_t)(uintptr_t)kern, (uint64_t)threads, (uint64_t)maxiters,
        (uint64_t)(uintptr_t)rank, (uint64_t)(uintptr_t)contrib[0], (uint64_t)(uintptr_t)contrib[1],
        (uint64_t)(uintptr_t)unit_part, (uint64_t)(uintptr_t)partial, (uint64_t)nparts,
        (uint64_t)(uintptr_t)g->off, (uint64_t)n, (uint64_t)(uintptr_t)done, (uint64_t)(uintptr_t)us.order,
        (uint64_t)us.grid, (uint64_t)(uintptr_t)us.heads, (uint64_t)us.nxcd, (uint64_t)(uintptr_t)us.shares,
        (uint64_t)(uintptr_t)B->pk, (uint64_t)(uintptr_t)B->seg_rel, (uint64_t)(uintptr_t)B->seg_hi,
        (uint64_t)(uintptr_t)B->unit_seg0, (uint64_t)B->sbits};
    GraphCache& gc = ctx->pr_graph;
    if (gc.key != key || !gc.exec) {
      gc.release();
      RG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int rc = iterations(false);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != RG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      RG_HIP(ec);
      gc.graph = graph;
      RG_HIP(hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0));
      gc.key = key;
    }
    RG_HIP(hipGraphLaunch(gc.exec, ctx->stream));
  } else {
    RG_TRY(iterations(true));
  }
  RG_HIP(hipEventRecord(t1, ctx->stream));
  RG_HIP(hipEventSynchronize(t1));
  float ms = 0.f;
  RG_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  int64_t dn = 0;
  RG_TRY(ctx_read_count(ctx, done, &dn, 1));
  int64_t last = dn ? dn : maxiters;
  double l1 = 0.0;
  if (last >= 1) RG_TRY(ctx_copy(ctx, &l1, l1_hist + (last - 1), sizeof(double), hipMemcpyDeviceToHost));
  if (iters_out) *iters_out = dn ? dn : maxiters + 1;
  if (l1_out) *l1_out = l1;
  if (ms_out) *ms_out = ms;
  ctx->last_result = rank;
  if (ranks_out) RG_TRY(ctx_copy(ctx, ranks_out, rank, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

}  // namespace rg
