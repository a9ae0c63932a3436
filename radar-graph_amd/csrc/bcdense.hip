// Dense BC levels on the destination-blocked layout (blocked.h): the
// forward path counting of BC_F (push-new-orig Ligra/apps/BC-manyiter-v2.C:
// 189-213, NumPaths[d] += NumPaths[s] for unvisited d) and the backward
// accumulation of BC_Back_F (215-236, Dependencies[d] += Dependencies[s] on
// the transposed graph), for frontiers large enough that the push engine
// would scatter an FP64 atomic per arc.  Those atomics execute at the memory
// side on this part; here one workgroup per unit sums the values of the
// frontier sources of its block's in-edges in LDS and writes each
// destination once.  Blocks split over several units (hub blocks) write
// per-unit partial sums that a second kernel adds in unit order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "blocked.h"
#include "device.h"
#include "engine.h"
#include "engine_host.h"
#include "queries.h"

namespace rg {

namespace {

constexpr int kBcThreads = 512;

struct BcDenseArgs {
  int64_t n;
  const uint64_t* off;                 // out-offsets of the orientation (frontier degrees)
  const uint32_t* src;
  const uint16_t* vl;
  const uint64_t* unit_lo;
  const uint32_t* unit_block;
  const uint32_t* block_unit0;
  const uint32_t* multi;
  const uint32_t* mchunk;              // first unit of each merge chunk (BlockedCSC::mchunk)
  const unsigned long long* active;    // bit per vertex: in the frontier
  const uint8_t* visited;
  const double* val;                   // value pushed from each frontier source
  double* out;                         // forward: set at joining vertices; backward: added
  double* part;                        // per-unit partial sums of split blocks
  uint32_t* out_ids;                   // forward: next frontier
  int64_t* out_info;                   // forward: [count, out-degree sum]
  unsigned long long* next_bits;       // forward (nullable): the next frontier's bitmap, a word per
                                       // wave of finalised destinations (blocks left out stay as they are)
};

// Whether every bit of the block's skip words is set (all threads read the
// same LDS words, so the answer is uniform across the workgroup).
template <int BV>
__device__ __forceinline__ bool block_done(const unsigned long long* skip) {
  bool all = true;
#pragma unroll 4
  for (int i = 0; i < BV / 64; ++i) all = all && skip[i] == ~0ull;
  return all;
}

// Destination d of the block, with sum x: forward, d joins the next
// frontier when a frontier in-neighbour reached it (x != 0: path counts of
// visited vertices are positive) and NumPaths[d] (0 while unvisited) = x;
// backward, Dependencies[d] += x.  bc_store writes the values and the next
// frontier's bit word and returns whether d joins; bc_finalize also queues
// the joining vertices on the wave's queue.
template <bool FWD>
__device__ __forceinline__ bool bc_store(const BcDenseArgs& a, int64_t d, bool valid, double x) {
  bool join = false;
  if (valid && x != 0.0) {
    if (FWD) {
      a.out[d] = x;
      join = true;
    } else {
      a.out[d] += x;
    }
  }
  if (FWD && a.next_bits) {
    // Called by whole waves over 64 aligned destinations.
    const unsigned long long word = __ballot(join);
    if (__lane_id() == 0 && d < a.n) a.next_bits[d >> 6] = word;
  }
  return join;
}

template <bool FWD>
__device__ __forceinline__ void bc_finalize(const BcDenseArgs& a, WaveQueue& q, const EdgeArgs& ea, int64_t d,
                                            bool valid, double x) {
  const bool join = bc_store<FWD>(a, d, valid, x);
  if (FWD) wq_push(q, ea, join, (uint32_t)d);
}

// kDenseArcs arcs of a unit per thread and step [e0 + j * THREADS]: their
// loads (arc, then the source's frontier bit word) issued together before
// any is tested.  ok[j]: the arc exists, its destination is not skipped and
// its source is on the frontier.
constexpr int kDenseArcs = 4;

template <int THREADS>
__device__ __forceinline__ void load_arcs(const BcDenseArgs& a, uint64_t e0, uint64_t hi,
                                          const unsigned long long* skip, uint32_t* s, uint32_t* l, bool* ok) {
#pragma unroll
  for (int j = 0; j < kDenseArcs; ++j) {
    const uint64_t e = e0 + (uint64_t)j * THREADS;
    ok[j] = e < hi;
    s[j] = ok[j] ? __builtin_nontemporal_load(a.src + e) : 0u;
    l[j] = ok[j] ? (uint32_t)__builtin_nontemporal_load(a.vl + e) : 0u;
  }
  unsigned long long act[kDenseArcs];
#pragma unroll
  for (int j = 0; j < kDenseArcs; ++j) {
    ok[j] = ok[j] && !((skip[l[j] >> 6] >> (l[j] & 63)) & 1);
    act[j] = ok[j] ? a.active[s[j] >> 6] : 0ull;
  }
#pragma unroll
  for (int j = 0; j < kDenseArcs; ++j) ok[j] = ok[j] && ((act[j] >> (s[j] & 63)) & 1);
}

template <bool FWD, int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bc_dense(BcDenseArgs a) {
  constexpr int BV = 1 << VB;
  constexpr int NW = THREADS / kWave;
  __shared__ double acc[BV];
  __shared__ unsigned long long skip[BV / 64];
  __shared__ uint32_t join_ids[FWD ? BV : 1];   // forward: the block's joining vertices
  __shared__ uint32_t join_n;
  const WgList joins{join_ids, &join_n, FWD ? (uint32_t)BV : 1u};
  const int64_t u = blockIdx.x;
  const uint32_t b = a.unit_block[u];
  const int64_t base = (int64_t)b << VB;
  const bool single = a.block_unit0[b + 1] - a.block_unit0[b] == 1;
  if (threadIdx.x == 0) join_n = 0;
  for (int l = threadIdx.x; l < BV; l += THREADS) {
    acc[l] = 0.0;
    const int64_t d = base + l;
    const bool sk = d >= a.n || a.visited[d] != 0;
    const unsigned long long w = __ballot(sk);
    if (__lane_id() == 0) skip[l >> 6] = w;
  }
  __syncthreads();
  // A block whose destinations are all visited has nothing to do (the merge
  // of a split block reads no partial of a visited destination).
  if (block_done<BV>(skip)) return;
  const uint64_t lo = a.unit_lo[u], hi = a.unit_lo[u + 1];
  for (uint64_t e0 = lo + threadIdx.x; e0 < hi; e0 += (uint64_t)THREADS * kDenseArcs) {
    uint32_t s[kDenseArcs], l[kDenseArcs];
    bool ok[kDenseArcs];
    load_arcs<THREADS>(a, e0, hi, skip, s, l, ok);
    // Each value is loaded next to its add.  The values read (frontier
    // sources, all visited) and the destinations written (unvisited) are
    // disjoint sets, so no workgroup reads a value another one writes.  A
    // form that issued the four value loads ahead of the four adds once gave
    // run-to-run differences on a 2^10-vertex R-MAT and is not used.
#pragma unroll
    for (int j = 0; j < kDenseArcs; ++j)
      if (ok[j]) atomicAdd(&acc[l[j]], a.val[s[j]]);
  }
  __syncthreads();
  if (!single) {
    double* p = a.part + ((size_t)u << VB);
    for (int l = threadIdx.x; l < BV; l += THREADS) p[l] = acc[l];
    return;
  }
  EdgeArgs ea{};
  ea.off = a.off;
  ea.out_ids = a.out_ids;
  ea.out_info = a.out_info;
  uint64_t deg = 0;
  for (int l0 = 0; l0 < BV; l0 += THREADS) {
    const int l = l0 + threadIdx.x;
    const int64_t d = base + l;
    const bool join = bc_store<FWD>(a, d, !((skip[l >> 6] >> (l & 63)) & 1), acc[l]);
    if (FWD) wgl_push(joins, ea, join, (uint32_t)d, deg);
  }
  if (FWD) wgl_flush<THREADS>(joins, ea, deg);
}

// Split blocks are merged in two stages, so that a hub block (hundreds of
// units) is read by workgroups across the whole grid rather than by the few
// that finish it: first each merge chunk (up to kMergeChunk consecutive
// units of one block, BlockedCSC::mchunk) combines its units' partials into
// its first unit's slot, in unit order; then each vertex of the block
// combines its chunk heads, in order.  Only destinations that take the
// result are read and written.

// Sum of vertex l's partials over units u0, u0 + step, ... below u1: unit
// u0 + i * step into accumulator i mod K (K loads in flight), the K sums
// added pairwise.
template <int VB, int K>
__device__ __forceinline__ double part_sum(const double* part, uint32_t u0, uint32_t u1, uint32_t step, int l) {
  double t[K];
#pragma unroll
  for (int k = 0; k < K; ++k) t[k] = 0.0;
  uint32_t u = u0;
  for (; u + (K - 1) * step < u1; u += K * step) {
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] += part[((size_t)(u + k * step) << VB) + l];
  }
  for (int k = 0; u < u1; u += step, ++k) t[k] += part[((size_t)u << VB) + l];
#pragma unroll
  for (int w = K / 2; w >= 1; w /= 2)
#pragma unroll
    for (int k = 0; k < w; ++k) t[k] += t[k + w];
  return t[0];
}

// First merge stage of a dense BC level.  Grid: nmchunk * BV/THREADS.
template <int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bc_merge_chunk(BcDenseArgs a) {
  constexpr int PER = (1 << VB) / THREADS;
  const uint32_t u = a.mchunk[blockIdx.x / PER];
  const uint32_t b = a.unit_block[u];
  const int l = (int)(blockIdx.x % PER) * THREADS + threadIdx.x;
  const int64_t d = ((int64_t)b << VB) + l;
  if (d >= a.n || a.visited[d] != 0) return;
  const uint32_t u1 = a.block_unit0[b + 1];
  const uint32_t ue = u1 - u > kMergeChunk ? u + kMergeChunk : u1;
  a.part[((size_t)u << VB) + l] = part_sum<VB, (int)kMergeChunk>(a.part, u, ue, 1, l);
}

// Second stage: the chunk heads' sums added, then the destination is
// finalised.  Grid: nmulti * BV/THREADS.
template <bool FWD, int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bc_merge(BcDenseArgs a) {
  constexpr int BV = 1 << VB;
  constexpr int PER = BV / THREADS;
  const uint32_t b = a.multi[blockIdx.x / PER];
  const int l = (int)(blockIdx.x % PER) * THREADS + threadIdx.x;
  const int64_t d = ((int64_t)b << VB) + l;
  const bool valid = d < a.n && a.visited[d] == 0;
  const double x = valid ? part_sum<VB, 8>(a.part, a.block_unit0[b], a.block_unit0[b + 1], kMergeChunk, l) : 0.0;
  const bool join = bc_store<FWD>(a, d, valid, x);
  if (FWD) {   // the joins appended once per workgroup
    EdgeArgs ea{};
    ea.off = a.off;
    ea.out_ids = a.out_ids;
    ea.out_info = a.out_info;
    wg_push<THREADS>(ea, join, (uint32_t)d);
  }
}

// BFS_F (BFS.C:28-49) on a dense level: every unvisited destination of
// the block takes a frontier in-neighbour as its parent (the smallest-id
// one seen by the unit, any being a valid BFS parent); split blocks keep
// the first unit's candidate.  `out` is Parents (uint32).
template <int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bfs_dense(BcDenseArgs a, uint32_t* parents, uint32_t* part) {
  constexpr int BV = 1 << VB;
  constexpr int NW = THREADS / kWave;
  __shared__ uint32_t cand[BV];
  __shared__ unsigned long long skip[BV / 64];
  __shared__ uint32_t join_ids[BV];
  __shared__ uint32_t join_n;
  const WgList joins{join_ids, &join_n, (uint32_t)BV};
  const int64_t u = blockIdx.x;
  const uint32_t b = a.unit_block[u];
  const int64_t base = (int64_t)b << VB;
  const bool single = a.block_unit0[b + 1] - a.block_unit0[b] == 1;
  if (threadIdx.x == 0) join_n = 0;
  for (int l = threadIdx.x; l < BV; l += THREADS) {
    cand[l] = kNone;
    const int64_t d = base + l;
    const bool sk = d >= a.n || parents[d] != kNone;
    const unsigned long long w = __ballot(sk);
    if (__lane_id() == 0) skip[l >> 6] = w;
  }
  __syncthreads();
  if (block_done<BV>(skip)) return;
  const uint64_t lo = a.unit_lo[u], hi = a.unit_lo[u + 1];
  for (uint64_t e0 = lo + threadIdx.x; e0 < hi; e0 += (uint64_t)THREADS * kDenseArcs) {
    uint32_t s[kDenseArcs], l[kDenseArcs];
    bool ok[kDenseArcs];
    load_arcs<THREADS>(a, e0, hi, skip, s, l, ok);
#pragma unroll
    for (int j = 0; j < kDenseArcs; ++j)
      if (ok[j] && cand[l[j]] > s[j]) atomicMin(&cand[l[j]], s[j]);
  }
  __syncthreads();
  if (!single) {
    uint32_t* p = part + ((size_t)u << VB);
    for (int l = threadIdx.x; l < BV; l += THREADS) p[l] = cand[l];
    return;
  }
  EdgeArgs ea{};
  ea.off = a.off;
  ea.out_ids = a.out_ids;
  ea.out_info = a.out_info;
  uint64_t deg = 0;
  for (int l0 = 0; l0 < BV; l0 += THREADS) {
    const int l = l0 + threadIdx.x;
    const uint32_t c = cand[l];
    const bool join = c != kNone;
    if (join) parents[base + l] = c;
    wgl_push(joins, ea, join, (uint32_t)(base + l), deg);
  }
  wgl_flush<THREADS>(joins, ea, deg);
}

// First candidate (in unit order) of vertex l over units u0, u0 + step, ...
// below u1, eight loads at a time; kNone if none has one.
template <int VB>
__device__ __forceinline__ uint32_t first_candidate(const uint32_t* part, uint32_t u0, uint32_t u1, uint32_t step,
                                                    int l) {
  uint32_t c = kNone;
  for (uint32_t u = u0; u < u1 && c == kNone; u += 8 * step) {
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = u + k * step < u1 ? part[((size_t)(u + k * step) << VB) + l] : kNone;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c == kNone) c = t[k];
  }
  return c;
}

// First merge stage of a dense BFS level (as k_bc_merge_chunk: the chunk's
// first candidate into its head's slot).  Grid: nmchunk * BV/THREADS.
template <int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bfs_merge_chunk(BcDenseArgs a, const uint32_t* parents, uint32_t* part) {
  constexpr int PER = (1 << VB) / THREADS;
  const uint32_t u = a.mchunk[blockIdx.x / PER];
  const uint32_t b = a.unit_block[u];
  const int l = (int)(blockIdx.x % PER) * THREADS + threadIdx.x;
  const int64_t d = ((int64_t)b << VB) + l;
  if (d >= a.n || parents[d] != kNone) return;
  const uint32_t u1 = a.block_unit0[b + 1];
  part[((size_t)u << VB) + l] = first_candidate<VB>(part, u, u1 - u > kMergeChunk ? u + kMergeChunk : u1, 1, l);
}

// Second stage: the first candidate over the chunk heads.
template <int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_bfs_merge(BcDenseArgs a, uint32_t* parents, const uint32_t* part) {
  constexpr int BV = 1 << VB;
  constexpr int PER = BV / THREADS;
  const uint32_t b = a.multi[blockIdx.x / PER];
  const int l = (int)(blockIdx.x % PER) * THREADS + threadIdx.x;
  const int64_t d = ((int64_t)b << VB) + l;
  const uint32_t c = d < a.n && parents[d] == kNone
                         ? first_candidate<VB>(part, a.block_unit0[b], a.block_unit0[b + 1], kMergeChunk, l)
                         : kNone;
  const bool join = c != kNone;
  if (join) parents[d] = c;
  EdgeArgs ea{};
  ea.off = a.off;
  ea.out_ids = a.out_ids;
  ea.out_info = a.out_info;
  wg_push<THREADS>(ea, join, (uint32_t)d);   // joins appended once per workgroup
}

// Bottom-up BFS level: the push-pull tree's edgeMapDense (push-pull
// Ligra/ligra/ligra.h:62-88) with decodeInNghBreakEarly (vertex.h:13-27):
// every unvisited vertex scans its in-neighbours in order and takes the
// first one in the frontier as parent, ending the scan (BFS_F's cond turns
// false).  Light vertices: one thread each; in-degree above heavy_deg: one
// wavefront each (second kernel), 64 in-neighbours per step, the lowest
// lane that finds one wins -- the same parent the ordered scan picks.
__global__ void k_last_in_edges(const uint64_t* in_off, int64_t n, unsigned long long* end) {
  unsigned long long best = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    if (in_off[v + 1] > in_off[v]) best = (unsigned long long)v + 1;
  for (int o = 32; o > 0; o >>= 1) best = max(best, (unsigned long long)__shfl_xor(best, o));
  if (__lane_id() == 0 && best) atomicMax(end, best);
}

// 1 + the last vertex with in-edges in g's current orientation (cached).
__global__ void k_first_in(const uint64_t* in_off, const uint32_t* in_nbr, int64_t n, uint32_t* first) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t b = in_off[v];
    first[v] = in_off[v + 1] > b ? in_nbr[b] : kNone;
  }
}

}  // namespace

// The first in-neighbour of every vertex in the current orientation, built
// once per orientation (a view shares its base's array when the base has
// one for the same in-edges).
int in_first_neighbours(rg_graph* g, const uint32_t** out) {
  if (g->first_for != g->in_nbr || !g->first_in) {
    const rg_graph* base = g->view ? g->base : nullptr;
    const bool shared = base && g->first_in && g->first_in == base->first_in;
    if (g->first_in && !shared) (void)hipFree(g->first_in);
    g->first_in = nullptr;
    g->first_for = nullptr;
    if (base && base->first_in && base->first_for == g->in_nbr) {
      g->first_in = base->first_in;
    } else {
      rg_ctx* ctx = g->ctx;
      RG_HIP(hipMalloc(&g->first_in, sizeof(uint32_t) * std::max<int64_t>(g->n, 1)));
      hipLaunchKernelGGL(k_first_in, grid_for(ctx, g->n), 256, 0, ctx->stream, g->in_off, g->in_nbr, g->n,
                         g->first_in);
      RG_HIP(hipGetLastError());
    }
    g->first_for = g->in_nbr;
  }
  *out = g->first_in;
  return RG_OK;
}

namespace {

int in_edges_end(rg_graph* g, int64_t* end) {
  if (g->nz_for != g->in_off) {
    rg_ctx* ctx = g->ctx;
    int64_t* d = ctx->counts.as<int64_t>() + 19;
    RG_TRY(fill<int64_t>(ctx, d, 1, 0));
    hipLaunchKernelGGL(k_last_in_edges, grid_for(ctx, g->n), 256, 0, ctx->stream, g->in_off, g->n,
                       reinterpret_cast<unsigned long long*>(d));
    RG_HIP(hipGetLastError());
    RG_TRY(ctx_read_count(ctx, d, &g->nz_end, 1));
    g->nz_for = g->in_off;
  }
  *end = g->nz_end;
  return RG_OK;
}

// Thread per vertex (a wave holds 64 aligned vertices, kPullU groups of
// them per step): an unvisited light vertex takes its first in-neighbour on
// the frontier as parent.  The kPullU vertices' loads -- parent, offsets,
// first in-neighbour (from the per-vertex copy first_in, read coalesced
// rather than one in-list line per vertex) and its frontier bit -- are
// issued together; a vertex
// whose first in-neighbour is not on the frontier then scans on alone,
// kPullScan in-neighbours and their frontier words loaded per step.
// (Handing long scans to groups of lanes measured slower at every split
// point: fewer scans in flight.)  Every word of the next frontier's bitmap
// `next` below vertex `nn` is written (a ballot per wave), so the next pull
// level reads it as it is; vertices at or above nn have no in-edges.  Heavy
// vertices are queued for k_bfs_pull_heavy, which ORs in their bits.
constexpr int kPullU = 4;
constexpr int kPullScan = 4;   // in-neighbours tested per step of a long scan

// Picks the level's frontier and next bitmaps of a device-controlled pull
// level (PullSel::mode set): false when this level is not a pull level.
__device__ __forceinline__ bool pull_select(const PullSel& s, const unsigned long long*& active,
                                            unsigned long long*& next) {
  if (!s.mode) return true;
  if (*s.mode != s.want) return false;
  const int64_t c = *s.chain;
  active = c == 0 ? s.zero_bits : c == 1 ? s.full0 : s.full1;
  next = c == 1 ? s.full1 : s.full0;
  return true;
}

// Wave queue entries of the light pull kernel: a big bottom-up level joins
// tens of millions of vertices, and every queue flush is an atomic on the
// one count of the next-frontier list (same-address atomics serialise; with
// 256-entry queues the RMAT-26 query spent a third of its time on them).
constexpr int kPullQueue = 1024;

__global__ __launch_bounds__(256) void k_bfs_pull_light(const uint64_t* in_off, const uint32_t* in_nbr,
                                                        const uint32_t* first_in, const uint64_t* off, int64_t nn,
                                                        const unsigned long long* active, uint32_t* parents,
                                                        uint32_t* heavy, int64_t* heavy_count, int heavy_deg,
                                                        uint32_t* out_ids, int64_t* out_info,
                                                        unsigned long long* next, PullSel sel) {
  __shared__ uint32_t s_q[4][kPullQueue];
  if (!pull_select(sel, active, next)) return;
  EdgeArgs ea{};
  ea.off = off;
  ea.out_ids = out_ids;
  ea.out_info = out_info;
  WaveQueue q{s_q[threadIdx.x / kWave], 0, 0, kPullQueue};
  const int64_t step = (int64_t)gridDim.x * blockDim.x * kPullU;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x * kPullU; base < nn; base += step) {
    int64_t v[kPullU];
    bool cand[kPullU], join[kPullU];
    uint64_t b[kPullU], e[kPullU];
    uint32_t s0[kPullU];
#pragma unroll
    for (int k = 0; k < kPullU; ++k) {
      v[k] = base + (int64_t)k * blockDim.x + threadIdx.x;
      cand[k] = v[k] < nn && parents[v[k]] == kNone;
      join[k] = false;
    }
#pragma unroll
    for (int k = 0; k < kPullU; ++k) {
      b[k] = cand[k] ? in_off[v[k]] : 0;
      e[k] = cand[k] ? in_off[v[k] + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPullU; ++k) {
      // Heavy candidates (hubs, consecutive ids after a degree sort) go to
      // the wave-per-vertex kernel's list, one atomic per wave.
      const bool hv = cand[k] && e[k] - b[k] > (uint64_t)heavy_deg;
      const uint64_t hm = __ballot(hv);
      if (hm) {
        unsigned long long hb = 0;
        if (__lane_id() == 0) hb = atomicAdd((unsigned long long*)heavy_count, (unsigned long long)__popcll(hm));
        hb = __shfl(hb, 0);
        if (hv) heavy[hb + __popcll(hm & lanemask_lt())] = (uint32_t)v[k];
      }
      cand[k] = cand[k] && !hv && e[k] > b[k];
      s0[k] = cand[k] ? first_in[v[k]] : 0u;   // = in_nbr[b[k]]
    }
#pragma unroll
    for (int k = 0; k < kPullU; ++k) {
      if (cand[k] && ((active[s0[k] >> 6] >> (s0[k] & 63)) & 1)) {
        parents[v[k]] = s0[k];
        join[k] = true;
      }
    }
#pragma unroll
    for (int k = 0; k < kPullU; ++k) {
      if (cand[k] && !join[k]) {
        // kPullScan in-neighbours per step, their loads issued together;
        // the first of them on the frontier (in list order) is the parent.
        for (uint64_t j = b[k] + 1; j < e[k]; j += kPullScan) {
          uint32_t s[kPullScan];
          unsigned long long w[kPullScan];
#pragma unroll
          for (int t = 0; t < kPullScan; ++t) s[t] = j + t < e[k] ? in_nbr[j + t] : 0u;
#pragma unroll
          for (int t = 0; t < kPullScan; ++t) w[t] = j + t < e[k] ? active[s[t] >> 6] : 0ull;
          int hit = kPullScan;
#pragma unroll
          for (int t = kPullScan - 1; t >= 0; --t)
            if ((w[t] >> (s[t] & 63)) & 1) hit = t;
          if (hit < kPullScan) {
            parents[v[k]] = s[hit];
            join[k] = true;
            break;
          }
        }
      }
      const unsigned long long word = __ballot(join[k]);
      if (__lane_id() == 0 && v[k] < nn) next[v[k] >> 6] = word;
      wq_push(q, ea, join[k], (uint32_t)v[k]);
    }
  }
  wq_finish(q, ea);
}

__global__ __launch_bounds__(256) void k_bfs_pull_heavy(const uint64_t* in_off, const uint32_t* in_nbr,
                                                        const uint64_t* off, const unsigned long long* active,
                                                        uint32_t* parents, const uint32_t* heavy,
                                                        const int64_t* heavy_count, uint32_t* out_ids,
                                                        int64_t* out_info, unsigned long long* next, PullSel sel) {
  __shared__ uint32_t s_q[4][kQueue];
  if (!pull_select(sel, active, next)) return;
  EdgeArgs ea{};
  ea.off = off;
  ea.out_ids = out_ids;
  ea.out_info = out_info;
  WaveQueue q{s_q[threadIdx.x / kWave], 0, 0};
  const int lane = (int)__lane_id();
  const int64_t items = *heavy_count;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t it = wave; it < items; it += nwaves) {
    const uint32_t v = heavy[it];
    const uint64_t b = in_off[v], e = in_off[v + 1];
    bool join = false;
    for (uint64_t j0 = b; j0 < e; j0 += kWave) {
      const uint64_t j = j0 + lane;
      uint32_t s = 0;
      bool f = false;
      if (j < e) {
        s = in_nbr[j];
        f = (active[s >> 6] >> (s & 63)) & 1;
      }
      const unsigned long long m = __ballot(f);
      if (m) {
        const int first = __ffsll((long long)m) - 1;
        const uint32_t p = __shfl(s, first);
        if (lane == 0) {
          parents[v] = p;
          atomicOr(next + (v >> 6), 1ull << (v & 63));
        }
        join = true;
        break;
      }
    }
    wq_push(q, ea, join && lane == 0, v);
  }
  wq_finish(q, ea);
}

// A device-controlled pull level whose frontier is only a list (chain 0):
// its bits set into the zero bitmap before the level, cleared after.
__global__ void k_list_bits_sel(const uint32_t* ids, const int64_t* count, PullSel sel, bool set) {
  if (*sel.mode != sel.want || *sel.chain != 0) return;
  unsigned long long* bits = const_cast<unsigned long long*>(sel.zero_bits);
  const int64_t c = *count;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < c; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t v = ids[i];
    if (set) atomicOr(bits + (v >> 6), 1ull << (v & 63));
    else bits[v >> 6] = 0ull;   // every bit set in the word is a frontier vertex's
  }
}

__global__ void k_set_bits(const uint32_t* ids, int64_t count, unsigned long long* bits) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t v = ids[i];
    atomicOr(bits + (v >> 6), 1ull << (v & 63));
  }
}

__global__ void k_clear_bits(const uint32_t* ids, int64_t count, unsigned long long* bits) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    bits[ids[i] >> 6] = 0;
}

}  // namespace

// One dense level over the frontier ids[0, count) of g's current
// orientation: kind 0 = BC forward, 1 = BC backward, 2 = BFS.  Forward and
// BFS: out_ids / out_info receive the next frontier (the caller zeroes
// out_info).  The frontier bitmap and partial sums live in the
// context's bc_dense buffer; the bitmap is all zero between calls.
namespace {

// The context's frontier bitmap for an n-vertex graph followed by
// part_bytes of scratch; the bitmap is all zero on return.
int frontier_bitmap(rg_ctx* ctx, int64_t n, size_t part_bytes, unsigned long long** bits, void** part) {
  const int64_t nwords = n / 64 + 1;
  const size_t bits_bytes = ((sizeof(unsigned long long) * nwords + 255) / 256) * 256;
  DevBuf& sb = ctx->bc_dense;
  if (sb.bytes < bits_bytes + part_bytes) {
    sb.release();
    RG_TRY(sb.ensure(bits_bytes + part_bytes));
    ctx->bc_bits_zero = 0;
  }
  // The partials of a smaller graph may have overwritten words past its own
  // bitmap: clear the bitmap of this graph unless it is known to be zero.
  if (ctx->bc_bits_zero < nwords) RG_TRY(fill<unsigned long long>(ctx, sb.as<unsigned long long>(), nwords, 0ull));
  ctx->bc_bits_zero = nwords;
  *bits = sb.as<unsigned long long>();
  *part = sb.as<char>() + bits_bytes;
  return RG_OK;
}

int dense_level(rg_graph* g, int kind, const uint32_t* ids, int64_t count, const double* val, double* out,
                const uint8_t* visited, uint32_t* parents, uint32_t* out_ids, int64_t* out_info,
                const unsigned long long* active_bits = nullptr, unsigned long long* next_bits = nullptr) {
  rg_ctx* ctx = g->ctx;
  BlockedCSC* B = nullptr;
  RG_TRY(query_layout(g, &B));
  const size_t part_bytes = sizeof(double) * ((size_t)std::max<int64_t>(B->nunits, 1) << kQueryVB);
  unsigned long long* bits = nullptr;
  void* part = nullptr;
  RG_TRY(frontier_bitmap(ctx, g->n, part_bytes, &bits, &part));
  BcDenseArgs a{};
  a.n = g->n;
  a.off = g->off;
  a.src = B->src;
  a.vl = B->vl;
  a.unit_lo = B->unit_lo;
  a.unit_block = B->unit_block;
  a.block_unit0 = B->block_unit0;
  a.multi = B->multi;
  a.mchunk = B->mchunk;
  a.active = bits;
  a.visited = visited;
  a.val = val;
  a.out = out;
  a.part = static_cast<double*>(part);
  a.out_ids = out_ids;
  a.out_info = out_info;
  a.next_bits = next_bits;
  // The frontier: a complete bitmap the caller keeps, or the list's bits
  // set into the zero bitmap (cleared again below).
  if (active_bits) a.active = active_bits;
  const int grid = grid_for(ctx, count);
  if (!active_bits) hipLaunchKernelGGL(k_set_bits, grid, 256, 0, ctx->stream, ids, count, bits);
  constexpr int PER = (1 << kQueryVB) / kBcThreads;
  hipEvent_t* ev = timing_begin(ctx);
  if (kind == 2) {
    uint32_t* upart = reinterpret_cast<uint32_t*>(a.part);
    hipLaunchKernelGGL((k_bfs_dense<kQueryVB, kBcThreads>), (int)B->nunits, kBcThreads, 0, ctx->stream, a, parents,
                       upart);
    if (B->nmulti) {
      hipLaunchKernelGGL((k_bfs_merge_chunk<kQueryVB, kBcThreads>), (int)(B->nmchunk * PER), kBcThreads, 0,
                         ctx->stream, a, parents, upart);
      hipLaunchKernelGGL((k_bfs_merge<kQueryVB, kBcThreads>), (int)(B->nmulti * PER), kBcThreads, 0, ctx->stream,
                         a, parents, upart);
    }
  } else if (kind == 0) {
    hipLaunchKernelGGL((k_bc_dense<true, kQueryVB, kBcThreads>), (int)B->nunits, kBcThreads, 0, ctx->stream, a);
    if (B->nmulti) {
      hipLaunchKernelGGL((k_bc_merge_chunk<kQueryVB, kBcThreads>), (int)(B->nmchunk * PER), kBcThreads, 0,
                         ctx->stream, a);
      hipLaunchKernelGGL((k_bc_merge<true, kQueryVB, kBcThreads>), (int)(B->nmulti * PER), kBcThreads, 0,
                         ctx->stream, a);
    }
  } else {
    hipLaunchKernelGGL((k_bc_dense<false, kQueryVB, kBcThreads>), (int)B->nunits, kBcThreads, 0, ctx->stream, a);
    if (B->nmulti) {
      hipLaunchKernelGGL((k_bc_merge_chunk<kQueryVB, kBcThreads>), (int)(B->nmchunk * PER), kBcThreads, 0,
                         ctx->stream, a);
      hipLaunchKernelGGL((k_bc_merge<false, kQueryVB, kBcThreads>), (int)(B->nmulti * PER), kBcThreads, 0,
                         ctx->stream, a);
    }
  }
  if (ev) RG_HIP(hipEventRecord(ev[1], ctx->stream));
  if (!active_bits) hipLaunchKernelGGL(k_clear_bits, grid, 256, 0, ctx->stream, ids, count, bits);
  RG_HIP(hipGetLastError());
  return RG_OK;
}

}  // namespace

int bc_dense_level(rg_graph* g, bool fwd, const uint32_t* ids, int64_t count, const double* val, double* out,
                   const uint8_t* visited, uint32_t* out_ids, int64_t* out_info,
                   const unsigned long long* active_bits, unsigned long long* next_bits) {
  return dense_level(g, fwd ? 0 : 1, ids, count, val, out, visited, nullptr, out_ids, out_info, active_bits,
                     next_bits);
}

int set_frontier_bits(rg_ctx* ctx, const uint32_t* ids, int64_t count, unsigned long long* bits) {
  if (count <= 0) return RG_OK;
  hipLaunchKernelGGL(k_set_bits, grid_for(ctx, count), 256, 0, ctx->stream, ids, count, bits);
  RG_HIP(hipGetLastError());
  return RG_OK;
}

int bfs_dense_level(rg_graph* g, const uint32_t* ids, int64_t count, uint32_t* parents, uint32_t* out_ids,
                    int64_t* out_info) {
  return dense_level(g, 2, ids, count, nullptr, nullptr, nullptr, parents, out_ids, out_info);
}

int bfs_pull_buffers(rg_graph* g, int heavy_deg, PullBufs* pb) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (heavy_deg <= 0) heavy_deg = 256;
  // The same layout as bfs_pull_level's: the zero frontier bitmap, then two
  // complete bitmaps, the heavy counter and the heavy list.
  const int64_t nwords = n / 64 + 1;
  const size_t full_bytes = ((sizeof(unsigned long long) * nwords + 255) / 256) * 256;
  const size_t cap = (size_t)(g->m / heavy_deg + 1);
  unsigned long long* bits = nullptr;
  void* part = nullptr;
  RG_TRY(frontier_bitmap(ctx, n, 2 * full_bytes + sizeof(int64_t) * 2 + sizeof(uint32_t) * cap, &bits, &part));
  pb->bits = bits;
  pb->full[0] = static_cast<unsigned long long*>(part);
  pb->full[1] = reinterpret_cast<unsigned long long*>(static_cast<char*>(part) + full_bytes);
  pb->heavy = reinterpret_cast<uint32_t*>(reinterpret_cast<int64_t*>(static_cast<char*>(part) + 2 * full_bytes) + 2);
  pb->heavy_deg = heavy_deg;
  pb->nwords = nwords;
  RG_TRY(in_edges_end(g, &pb->nn));
  RG_TRY(in_first_neighbours(g, &pb->first_in));
  pb->wnz = (pb->nn + 63) / 64;
  return RG_OK;
}

int bfs_pull_zero_tails(rg_ctx* ctx, const PullBufs& pb) {
  if (pb.nwords > pb.wnz) {
    RG_TRY(fill<unsigned long long>(ctx, pb.full[0] + pb.wnz, pb.nwords - pb.wnz, 0ull));
    RG_TRY(fill<unsigned long long>(ctx, pb.full[1] + pb.wnz, pb.nwords - pb.wnz, 0ull));
  }
  return RG_OK;
}

int bfs_pull_level_dev(rg_graph* g, const PullBufs& pb, const PullSel& sel, const uint32_t* ids,
                       const int64_t* count_dev, uint32_t* parents, int64_t* hcount, uint32_t* out_ids,
                       int64_t* out_info) {
  rg_ctx* ctx = g->ctx;
  const int gl = ctx->num_cus * 4;
  hipLaunchKernelGGL(k_list_bits_sel, gl, 256, 0, ctx->stream, ids, count_dev, sel, true);
  const int vgrid = (int)std::min<int64_t>((pb.nn + 256 * kPullU - 1) / (256 * kPullU), (int64_t)ctx->num_cus * 8);
  hipLaunchKernelGGL(k_bfs_pull_light, std::max(vgrid, 1), 256, 0, ctx->stream, g->in_off, g->in_nbr, pb.first_in,
                     g->off, pb.nn, (const unsigned long long*)nullptr, parents, pb.heavy, hcount, pb.heavy_deg,
                     out_ids, out_info, (unsigned long long*)nullptr, sel);
  hipLaunchKernelGGL(k_bfs_pull_heavy, ctx->num_cus * 4, 256, 0, ctx->stream, g->in_off, g->in_nbr, g->off,
                     (const unsigned long long*)nullptr, parents, pb.heavy, hcount, out_ids, out_info,
                     (unsigned long long*)nullptr, sel);
  hipLaunchKernelGGL(k_list_bits_sel, gl, 256, 0, ctx->stream, ids, count_dev, sel, false);
  RG_HIP(hipGetLastError());
  return RG_OK;
}

int bfs_pull_level(rg_graph* g, const uint32_t* ids, int64_t count, uint32_t* parents, int heavy_deg,
                   uint32_t* out_ids, int64_t* out_info, const unsigned long long* cur_bits,
                   const unsigned long long** next_bits) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (heavy_deg <= 0) heavy_deg = 256;
  // After the frontier bitmap (all zero between uses): two complete
  // bitmaps that consecutive pull levels pass on, the heavy counter and the
  // heavy list (vertices above heavy_deg; at most m / heavy_deg).
  const int64_t nwords = n / 64 + 1;
  const size_t full_bytes = ((sizeof(unsigned long long) * nwords + 255) / 256) * 256;
  const size_t cap = (size_t)(g->m / heavy_deg + 1);
  unsigned long long* bits = nullptr;
  void* part = nullptr;
  RG_TRY(frontier_bitmap(ctx, n, 2 * full_bytes + sizeof(int64_t) * 2 + sizeof(uint32_t) * cap, &bits, &part));
  unsigned long long* full[2] = {static_cast<unsigned long long*>(part),
                                 reinterpret_cast<unsigned long long*>(static_cast<char*>(part) + full_bytes)};
  int64_t* hcount = reinterpret_cast<int64_t*>(static_cast<char*>(part) + 2 * full_bytes);
  uint32_t* heavy = reinterpret_cast<uint32_t*>(hcount + 2);
  RG_TRY(fill<int64_t>(ctx, hcount, 1, 0));
  // The current frontier: the previous pull level's complete bitmap, or
  // the list's bits set into the zero bitmap (cleared again below).
  const bool chained = cur_bits == full[0] || cur_bits == full[1];
  const unsigned long long* active = chained ? cur_bits : bits;
  unsigned long long* next = cur_bits == full[0] ? full[1] : full[0];
  // Vertices at or above nn have no in-edges: the pull kernel stops there
  // and their bitmap words stay zero (cleared when a chain starts).
  int64_t nn = n;
  RG_TRY(in_edges_end(g, &nn));
  const uint32_t* first_in = nullptr;
  RG_TRY(in_first_neighbours(g, &first_in));
  const int64_t wnz = (nn + 63) / 64;
  if (!chained && nwords > wnz) {
    RG_TRY(fill<unsigned long long>(ctx, full[0] + wnz, nwords - wnz, 0ull));
    RG_TRY(fill<unsigned long long>(ctx, full[1] + wnz, nwords - wnz, 0ull));
  }
  const int grid = grid_for(ctx, count);
  if (!chained) hipLaunchKernelGGL(k_set_bits, grid, 256, 0, ctx->stream, ids, count, bits);
  hipEvent_t* ev = timing_begin(ctx);
  const int vgrid = (int)std::min<int64_t>((nn + 256 * kPullU - 1) / (256 * kPullU), (int64_t)ctx->num_cus * 8);
  hipLaunchKernelGGL(k_bfs_pull_light, std::max(vgrid, 1), 256, 0, ctx->stream, g->in_off, g->in_nbr, first_in,
                     g->off, nn, active, parents, heavy, hcount, heavy_deg, out_ids, out_info, next, PullSel{});
  hipLaunchKernelGGL(k_bfs_pull_heavy, ctx->num_cus * 4, 256, 0, ctx->stream, g->in_off, g->in_nbr, g->off, active,
                     parents, heavy, hcount, out_ids, out_info, next, PullSel{});
  if (ev) RG_HIP(hipEventRecord(ev[1], ctx->stream));
  if (!chained) hipLaunchKernelGGL(k_clear_bits, grid, 256, 0, ctx->stream, ids, count, bits);
  RG_HIP(hipGetLastError());
  if (next_bits) *next_bits = next;
  return RG_OK;
}

}  // namespace rg
