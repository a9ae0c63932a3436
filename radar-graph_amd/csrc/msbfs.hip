// Bit-parallel multi-source BFS: up to 256 single-source BFS queries
// traversed together, source i owning bit i of W = 1, 2 or 4 64-bit words
// per vertex.  This is Radii's visited-word trick (push-new-orig
// Ligra/apps/Radii.C:35-75: NextVisited[d] |= Visited[s], one word, 64
// sources) generalised to several words and to per-source results, so a
// batch of BFS queries shares every edge scan.
//
// Per vertex: seen (bits of the sources that reached it), front (bits for
// which it is on the current level; two buffers, current and next) and, for
// sparse levels, next (bits arriving this level).  A level is either
//   * sparse: the push engine (engine.h) over the frontier list with an
//     operator that ORs front[s] & ~seen[d] into next[d] (global atomics,
//     few at this frontier size), then a vertex pass over the touched list;
//   * dense: the destination-blocked in-edge layout (blocked.h): one
//     workgroup per unit ORs front[s] of active sources into an LDS copy of
//     its block's words and finalises the block itself -- no global
//     atomics; blocks split over several units (hubs) are merged by a
//     second kernel.
// Ligra's rule picks the level kind (|F| + sum of out-degrees > threshold,
// ligra.h:269), on the device.  Two launch forms give the same answers:
//   * per-level kernels: every level launches the kernels of both kinds and
//     the ones not chosen exit at once; chunks of four levels are captured
//     once as a HIP graph and replayed, and the host reads the state back
//     once per chunk;
//   * one persistent launch: a cooperative grid runs every level of the
//     batch, whatever its source count (1..256), with grid barriers between
//     the phases of a level (k_ms_persist).
//
// Per source the batch yields the BFS summary of a single-source run:
// rounds, vertices reached, the out-degree sum of the reached vertices and
// a checksum of its level array (sum over reached v of mix(v, level) mod
// 2^64), and optionally the level arrays themselves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "bcbatch.h"
#include "blocked.h"
#include "device.h"
#include "engine.h"
#include "engine_host.h"
#include "queries.h"

namespace rg {

constexpr int kMsVB = kQueryVB;      // 4096 destinations per dense block
// Dense-level workgroup: the LDS copy of a 4096-vertex block (W words per
// vertex) allows 4 / 2 / 1 workgroups per CU for W = 1 / 2 / 4, so wider
// batches use wider workgroups to keep as many gathers in flight.
template <int W>
constexpr int ms_threads() { return W == 1 ? 512 : 1024; }
constexpr int kMsMaxSources = 256;
constexpr int kMsLevelsPerSync = 4;  // levels launched between host checks
// Dense level: arcs per thread and step.  Eight for one word per vertex
// (RMAT-26 batch of 64: 25.7 ms against 26.4 with four), four for two and
// four words (eight measured slower for two; four words already fill the
// register budget of a 1024-thread workgroup).
template <int W>
constexpr int ms_arcs() { return W == 1 ? 8 : 4; }

enum { MS_SPARSE = 0, MS_DENSE = 1, MS_DONE = 2 };
static_assert(MS_DONE == kMsDone, "level hooks test the mode against kMsDone");

// Device control block of one batch.
struct MsCtl {
  int64_t mode;
  int64_t level;          // level of the current frontier
  int64_t fcount, fdeg;   // current frontier list: size, out-degree sum
  int64_t ncount, ndeg;   // next frontier list (the engine's out_info)
  int64_t threshold;
  int64_t heavy_count;    // engine heavy queue
  int64_t dense_levels, sparse_levels;
  int64_t unit_next;      // dense level: next unit to take (dynamic work distribution)
  // Read by the level kernels on the device, so that a captured chunk of
  // levels is the same for every batch of one word width: the batch's
  // source count and the stamp tag of the current level.
  int64_t nsrc;
  uint32_t tag;
};

typedef unsigned long long ull;

struct MsArgs {
  MsCtl* ctl;
  int64_t n;
  int nsrc;
  const uint64_t* off;                 // out-offsets (degrees)
  // blocked in-edges
  const uint32_t* src;
  const uint16_t* vl;
  const uint64_t* unit_lo;
  const uint32_t* unit_block;
  const uint32_t* block_unit0;
  const uint32_t* multi;
  const uint32_t* mchunk;              // first unit of each merge chunk (BlockedCSC::mchunk)
  // per-vertex words
  ull* seen;
  const ull* front_cur;
  ull* front_nxt;
  ull* next;
  const ull* active_cur;               // bit per vertex: front_cur != 0
  ull* active_nxt;
  const uint32_t* list_cur;
  uint32_t* list_nxt;
  ull* part;                           // nunits << VB words * W (split blocks)
  // per source
  ull* lvl_new;
  ull* edges;
  ull* cs;
  int32_t* levels;                     // optional [nsrc][n]
  int32_t* radii;                      // optional [n]: last level a vertex gained a bit (Radii.C)
  uint32_t wide;                       // m >= 2^32: the statistics keep count and degree sum apart
};

// ---------------------------------------------------------------------------
// Sparse level: the push engine's operator.

template <int W>
struct MsWords {
  ull w[W];
};

template <int W>
struct MsPushOp {
  using SrcT = MsWords<W>;
  using HubT = uint32_t;
  static constexpr bool kOutput = true;
  static constexpr bool kL2 = false;
  const ull* front;
  const ull* seen;
  ull* next;
  uint32_t* stamp;
  const uint32_t* tag;   // the level's stamp tag (MsCtl::tag)
  const int64_t* mode;
  mutable uint32_t t = 0;   // *tag, read once per kernel by skip(), which every edge kernel calls first
  __device__ bool skip() const {
    t = *tag;
    return *mode != MS_SPARSE;
  }
  __device__ SrcT load_src(uint32_t s) const {
    SrcT r;
#pragma unroll
    for (int w = 0; w < W; ++w) r.w[w] = front[(size_t)s * W + w];
    return r;
  }
  __device__ bool cond(uint32_t) const { return true; }
  __device__ bool pre_hub(const SrcT&, uint32_t) const { return true; }
  // next[d] |= front[s] & ~seen[d]; d joins the touched list once per level.
  __device__ bool update(uint32_t, SrcT sv, uint32_t d) const {
    bool any = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ull x = sv.w[w] & ~seen[(size_t)d * W + w];
      if (x) {
        atomicOr(next + (size_t)d * W + w, x);
        any = true;
      }
    }
    if (!any || stamp[d] == t) return false;
    return atomicExch(stamp + d, t) != t;
  }
  __device__ uint32_t hub_identity() const { return 0; }
  __device__ void hub_update(uint32_t*, uint32_t, SrcT, uint32_t) const {}
  __device__ bool hub_flush(uint32_t, uint32_t) const { return false; }
  template <class V>
  __device__ void l2_update(uint32_t*, const V&) const {}
  __device__ uint32_t hub_combine(uint32_t x, uint32_t) const { return x; }
};

// ---------------------------------------------------------------------------
// allbits[w]: the bits of word w that hold one of the batch's nsrc sources.
template <int W>
__device__ __forceinline__ void batch_bits(int nsrc, ull* allbits) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int rem = nsrc - 64 * w;
    allbits[w] = rem >= 64 ? ~0ull : rem <= 0 ? 0ull : ((1ull << rem) - 1);
  }
}

// Per-source statistics staged in LDS, flushed once per workgroup.
//
// A vertex reached by source s adds 1 to s's count, its out-degree to s's
// edge sum and level_hash(v, level) to s's checksum.  With m < 2^32 the
// count and the edge sum share one 64-bit word, count << 32 | degree sum
// (a workgroup's degree sum for one source stays below m < 2^32, so the low
// half never carries): two LDS atomics per (vertex, source) pair instead of
// three.  A word of a vertex's new bits with more than 32 of its 64 sources
// set is recorded by its complement: the vertex is added once to the
// word's "every source" sums (one slot per lane, so the lanes of an atomic
// never collide) and subtracted for the sources it did not reach, which
// the flush adds back (mod 2^64, exact).  Graphs with m >= 2^32, and calls
// that write every level, walk the set bits with separate counters.
template <int W>
struct MsStats {
  uint32_t cnt[64 * W];   // m >= 2^32 only
  ull deg[64 * W];        // count << 32 | degree sum (m < 2^32), or the degree sum
  ull cs[64 * W];
  ull all_cd[W][64];      // per word and lane: sums added to every source of the word
  ull all_cs[W][64];
};

template <int W, int THREADS>
__device__ __forceinline__ void stats_init(MsStats<W>& st) {
  for (int i = threadIdx.x; i < 64 * W; i += THREADS) {
    st.cnt[i] = 0;
    st.deg[i] = 0;
    st.cs[i] = 0;
    st.all_cd[i >> 6][i & 63] = 0;
    st.all_cs[i >> 6][i & 63] = 0;
  }
}

__device__ __forceinline__ ull wave_sum_u64(ull v) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
  return v;
}

// Every thread of the workgroup calls it.
template <int W, int THREADS>
__device__ __forceinline__ void stats_flush(const MsArgs& a, MsStats<W>& st) {
  constexpr int NW = THREADS / kWave;
  __syncthreads();
  if (!a.wide) {
    // Fold each word's per-lane "every source" sums into slot 0.
    for (int p = (int)(threadIdx.x / kWave); p < 2 * W; p += NW) {
      ull* slot = (p & 1) ? st.all_cs[p >> 1] : st.all_cd[p >> 1];
      const ull v = wave_sum_u64(slot[__lane_id()]);
      if (__lane_id() == 0) slot[0] = v;
    }
    __syncthreads();
  }
  const int nsrc = (int)a.ctl->nsrc;
  for (int i = threadIdx.x; i < nsrc; i += THREADS) {
    ull cnt, deg, cs = st.cs[i];
    if (a.wide) {
      cnt = st.cnt[i];
      deg = st.deg[i];
    } else {
      const ull cd = st.deg[i] + st.all_cd[i >> 6][0];
      cs += st.all_cs[i >> 6][0];
      cnt = cd >> 32;
      deg = cd & 0xffffffffull;
    }
    if (cnt) {
      atomicAdd(a.lvl_new + i, cnt);
      atomicAdd(a.edges + i, deg);
      atomicAdd(a.cs + i, cs);
    }
  }
}

// Records vertex d, reached at `level` by the sources in nn (allbits: the
// batch's sources).  The lanes of a wave often hold the same sources (a
// hub block's vertices are reached by the same sources at the same level),
// so lane r walks its bits from bit r on, rotating: the lanes of one LDS
// atomic then mostly add into different sources' counters instead of
// queueing on one.
template <int W>
__device__ __forceinline__ void stats_add(const MsArgs& a, MsStats<W>& st, uint32_t d, uint32_t level,
                                          const ull* nn, ull deg, const ull* allbits) {
  const ull h = level_hash(d, level);
  const int r = (int)__lane_id();
  if (a.wide) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ull bits = __builtin_rotateright64(nn[w], (ull)r);
      while (bits) {
        const int s = w * 64 + ((__ffsll((long long)bits) - 1 + r) & 63);
        bits &= bits - 1;
        atomicAdd(&st.cnt[s], 1u);
        atomicAdd(&st.deg[s], deg);
        atomicAdd(&st.cs[s], h);
        if (a.levels) a.levels[(size_t)s * a.n + d] = (int32_t)level;
      }
    }
    return;
  }
  const ull cd = (1ull << 32) | deg;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    ull x = nn[w], vcd = cd, vcs = h;
    if (!a.levels && __popcll(x) > 32) {
      atomicAdd(&st.all_cd[w][r], cd);
      atomicAdd(&st.all_cs[w][r], h);
      x = ~x & allbits[w];
      vcd = 0ull - cd;
      vcs = 0ull - h;
    }
    ull bits = __builtin_rotateright64(x, (ull)r);
    while (bits) {
      const int s = w * 64 + ((__ffsll((long long)bits) - 1 + r) & 63);
      bits &= bits - 1;
      atomicAdd(&st.deg[s], vcd);
      atomicAdd(&st.cs[s], vcs);
      if (a.levels) a.levels[(size_t)s * a.n + d] = (int32_t)level;
    }
  }
}

// Sparse level, vertex pass over the touched list (= the next frontier):
// nn = next[d] (already disjoint from seen), seen |= nn, front_nxt = nn.
// Entries [first, count) with stride `step`.
template <int W>
__device__ __forceinline__ void ms_sparse_fin(const MsArgs& a, MsStats<W>& st, int64_t count, uint32_t level,
                                              int64_t first, int64_t step, const ull* allbits) {
  for (int64_t i = first; i < count; i += step) {
    const uint32_t d = a.list_nxt[i];
    ull nn[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const size_t k = (size_t)d * W + w;
      nn[w] = a.next[k];
      a.next[k] = 0;
      if (nn[w]) {
        a.seen[k] |= nn[w];
        a.front_nxt[k] = nn[w];
      }
    }
    atomicOr(a.active_nxt + (d >> 6), 1ull << (d & 63));
    if (a.radii) a.radii[d] = (int32_t)level;
    stats_add<W>(a, st, d, level, nn, a.off[d + 1] - a.off[d], allbits);
  }
}

template <int W>
__global__ __launch_bounds__(256) void k_ms_sparse_fin(MsArgs a) {
  __shared__ MsStats<W> st;
  if (a.ctl->mode != MS_SPARSE) return;
  ull allbits[W];
  batch_bits<W>((int)a.ctl->nsrc, allbits);
  stats_init<W, 256>(st);
  __syncthreads();
  ms_sparse_fin<W>(a, st, a.ctl->ncount, (uint32_t)a.ctl->level + 1, blockIdx.x * (int64_t)blockDim.x + threadIdx.x,
                   (int64_t)gridDim.x * blockDim.x, allbits);
  stats_flush<W, 256>(a, st);
}

// ---------------------------------------------------------------------------
// Dense level.

// Finalises destination d (valid lanes) from its OR-accumulated words:
// seen, front_nxt (every vertex of the block is written), the active bit
// word (lane 0 of each wave; the 64 lanes hold 64 aligned vertices), the
// next-frontier list and the statistics.  Called by whole waves.
template <int W>
__device__ __forceinline__ void ms_finalize(const MsArgs& a, MsStats<W>& st, const WgList& L, uint64_t& qdeg,
                                            const EdgeArgs& ea, int64_t d, bool valid, const ull* accw,
                                            uint32_t level, const ull* allbits) {
  ull nn[W];
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    nn[w] = 0;
    if (valid) {
      const size_t k = (size_t)d * W + w;
      if (accw[w]) {
        const ull sn = a.seen[k];
        nn[w] = accw[w] & ~sn & allbits[w];
        if (nn[w]) {
          a.front_nxt[k] = nn[w];
          a.seen[k] = sn | nn[w];
        }
      }
    }
    any |= nn[w] != 0;
  }
  // front_nxt and active_nxt are all zero when a level starts (each level
  // clears what it used), so only new words are written.
  const ull word = __ballot(any);
  if (__lane_id() == 0 && word) a.active_nxt[d >> 6] = word;
  ull deg = 0;
  if (any) {
    deg = a.off[d + 1] - a.off[d];
    if (a.radii) a.radii[d] = (int32_t)level;
    stats_add<W>(a, st, (uint32_t)d, level, nn, deg, allbits);
  }
  wgl_push(L, ea, any, (uint32_t)d, qdeg);
}

template <class T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Dense level over units [0, nunits) taken from the counter *unit_next:
// a persistent workgroup takes the next unit until none is left (the
// conditions below are uniform across it).  acc (BV * W words), full
// (BV / 64) and fullw (BV bytes when W > 1) are the workgroup's LDS.
template <int W, int VB, int THREADS>
__device__ __forceinline__ void ms_dense_units(const MsArgs& a, int64_t nunits, unsigned long long* unit_next,
                                               uint32_t level, const ull* allbits, ull* acc, ull* full,
                                               uint8_t* fullw, MsStats<W>& st, const WgList& L, uint64_t& qdeg,
                                               const EdgeArgs& ea, int64_t& s_unit) {
  constexpr int BV = 1 << VB;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_unit = (int64_t)atomicAdd(unit_next, 1ull);
    __syncthreads();
    const int64_t u = s_unit;
    if (u >= nunits) break;
    const uint32_t b = a.unit_block[u];
    const int64_t base = (int64_t)b << VB;
    const bool single = a.block_unit0[b + 1] - a.block_unit0[b] == 1;
    for (int i = threadIdx.x; i < BV * W; i += THREADS) acc[i] = 0;
    // Destinations every source of the batch has reached need no gathers;
    // with several words, neither do the words whose sources all reached
    // the destination (their bits would be masked by seen).
    for (int l = threadIdx.x; l < BV; l += THREADS) {
      const int64_t d = base + l;
      uint32_t fm = (1u << W) - 1;
      if (d < a.n) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          if ((a.seen[(size_t)d * W + w] & allbits[w]) != allbits[w]) fm &= ~(1u << w);
      }
      if (W > 1) fullw[l] = (uint8_t)fm;
      const ull word = __ballot(fm == (1u << W) - 1);
      if (__lane_id() == 0) full[l >> 6] = word;
    }
    __syncthreads();
    // Every destination reached by every source: nothing arrives (the merge
    // of a split block masks its partials with ~seen, which is then empty).
    bool all = true;
    for (int i = 0; i < BV / 64; ++i) all = all && full[i] == ~0ull;
    if (all) continue;
    // A = ms_arcs<W>() arcs per thread and step, each stage's loads issued for all
    // of them before any result is tested: the arc, its active bit word,
    // then the source's W frontier words (kept in flight together).
    constexpr int A = ms_arcs<W>();
    const uint64_t lo = a.unit_lo[u], hi = a.unit_lo[u + 1];
    for (uint64_t e0 = lo + threadIdx.x; e0 < hi; e0 += (uint64_t)THREADS * A) {
      uint32_t s[A], l[A];
      bool ok[A];
#pragma unroll
      for (int j = 0; j < A; ++j) {
        const uint64_t e = e0 + (uint64_t)j * THREADS;
        ok[j] = e < hi;
        s[j] = ok[j] ? ld_nt(a.src + e) : 0u;
        l[j] = ok[j] ? (uint32_t)ld_nt(a.vl + e) : 0u;
      }
      ull act[A];
      uint32_t fm[A];
#pragma unroll
      for (int j = 0; j < A; ++j) {
        ok[j] = ok[j] && !((full[l[j] >> 6] >> (l[j] & 63)) & 1);
        act[j] = ok[j] ? a.active_cur[s[j] >> 6] : 0ull;
        fm[j] = W > 1 && ok[j] ? fullw[l[j]] : 0u;
      }
      ull f[A][W];
#pragma unroll
      for (int j = 0; j < A; ++j) {
        ok[j] = ok[j] && ((act[j] >> (s[j] & 63)) & 1);
        const ull* p = a.front_cur + (size_t)s[j] * W;
#pragma unroll
        for (int w = 0; w < W; ++w) f[j][w] = ok[j] && !((fm[j] >> w) & 1) ? p[w] : 0ull;
      }
#pragma unroll
      for (int j = 0; j < A; ++j)
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (f[j][w]) atomicOr(&acc[l[j] * W + w], f[j][w]);
    }
    __syncthreads();
    if (!single) {
      ull* out = a.part + ((size_t)u << VB) * W;
      for (int i = threadIdx.x; i < BV * W; i += THREADS) out[i] = acc[i];
      continue;
    }
    for (int l0 = 0; l0 < BV; l0 += THREADS) {
      const int l = l0 + threadIdx.x;
      const int64_t d = base + l;
      ms_finalize<W>(a, st, L, qdeg, ea, d, d < a.n, &acc[l * W], level, allbits);
      wgl_drain<THREADS>(L, ea, THREADS);
    }
  }
}

template <int W, int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_ms_dense(MsArgs a, int64_t nunits) {
  constexpr int BV = 1 << VB;
  constexpr int NW = THREADS / kWave;
  __shared__ ull acc[BV * W];
  __shared__ ull full[BV / 64];
  __shared__ uint8_t fullw[W > 1 ? BV : 1];   // W > 1: bit w = every source of word w reached it
  __shared__ uint32_t join_ids[NW * kQueue];
  __shared__ uint32_t join_n;
  __shared__ MsStats<W> st;
  __shared__ int64_t s_unit;
  if (a.ctl->mode != MS_DENSE) return;
  const uint32_t level = (uint32_t)a.ctl->level + 1;
  ull allbits[W];
  batch_bits<W>((int)a.ctl->nsrc, allbits);
  EdgeArgs ea{};
  ea.off = a.off;
  ea.out_ids = a.list_nxt;
  ea.out_info = &a.ctl->ncount;
  const WgList joins{join_ids, &join_n, (uint32_t)(NW * kQueue)};
  if (threadIdx.x == 0) join_n = 0;
  uint64_t qdeg = 0;
  stats_init<W, THREADS>(st);
  ms_dense_units<W, VB, THREADS>(a, nunits, (unsigned long long*)&a.ctl->unit_next, level, allbits, acc, full, fullw,
                                 st, joins, qdeg, ea, s_unit);
  wgl_flush<THREADS>(joins, ea, qdeg);
  stats_flush<W, THREADS>(a, st);
}

// Blocks split over several units (the hub blocks of a degree-sorted
// graph, hundreds of units each) are merged in two stages, so that a hub
// block is read by many workgroups, not by the few that finalise it:
// first every merge chunk (up to kMergeChunk consecutive units of one
// block) ORs its units' partial words into the partial slot of its first
// unit; then each vertex ORs its block's chunk heads and is finalised.
// Only the words whose sources have not all reached the vertex are loaded
// and written (ms_finalize masks the others with ~seen, and seen does not
// change between the stages).  Loads of U units are issued before any is
// combined: 8 in the merge kernels; in the persistent kernel, whose
// registers are sized by its dense phase, 8 / W with W = 1 or 2 and 2 with
// W = 4 (more would spill there).

// Bit w set: word w of vertex d still misses some of the batch's sources.
template <int W>
__device__ __forceinline__ uint32_t ms_need(const MsArgs& a, int64_t d, const ull* allbits) {
  uint32_t need = 0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    if ((a.seen[(size_t)d * W + w] & allbits[w]) != allbits[w]) need |= 1u << w;
  return need;
}

// accw = OR of the partial words (need-masked) of vertex l over units
// u0, u0 + step, ... below u1.
template <int W, int VB, int U>
__device__ __forceinline__ void ms_or_units(const MsArgs& a, uint32_t u0, uint32_t u1, uint32_t step, int l,
                                            uint32_t need, ull* accw) {
#pragma unroll
  for (int w = 0; w < W; ++w) accw[w] = 0;
  uint32_t u = u0;
  for (; u + (U - 1) * step < u1; u += U * step) {
    ull t[U][W];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const ull* p = a.part + (((size_t)(u + k * step) << VB) + l) * W;
#pragma unroll
      for (int w = 0; w < W; ++w) t[k][w] = (need >> w) & 1 ? p[w] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
#pragma unroll
      for (int w = 0; w < W; ++w) accw[w] |= t[k][w];
  }
  for (; u < u1; u += step) {
    const ull* p = a.part + (((size_t)u << VB) + l) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) accw[w] |= (need >> w) & 1 ? p[w] : 0ull;
  }
}

// First stage, vertex l of the merge chunk that starts at unit u.
template <int W, int VB, int U>
__device__ __forceinline__ void ms_merge_chunk(const MsArgs& a, uint32_t u, int l, const ull* allbits) {
  const uint32_t b = a.unit_block[u];
  const int64_t d = ((int64_t)b << VB) + l;
  if (d >= a.n) return;
  const uint32_t need = ms_need<W>(a, d, allbits);
  if (!need) return;
  const uint32_t u1 = a.block_unit0[b + 1];
  ull accw[W];
  ms_or_units<W, VB, U>(a, u, u1 - u > kMergeChunk ? u + kMergeChunk : u1, 1, l, need, accw);
  ull* p = a.part + (((size_t)u << VB) + l) * W;
#pragma unroll
  for (int w = 0; w < W; ++w)
    if ((need >> w) & 1) p[w] = accw[w];
}

// Second stage: OR the chunk heads of split block b for vertex l and
// finalise it.  Called by whole waves (64 aligned vertices).
template <int W, int VB, int U>
__device__ __forceinline__ void ms_merge_vertex(const MsArgs& a, MsStats<W>& st, const WgList& L, uint64_t& qdeg,
                                                const EdgeArgs& ea, uint32_t b, int l, uint32_t level,
                                                const ull* allbits) {
  const int64_t d = ((int64_t)b << VB) + l;
  const bool valid = d < a.n;
  const uint32_t need = valid ? ms_need<W>(a, d, allbits) : 0u;
  ull accw[W];
  ms_or_units<W, VB, U>(a, a.block_unit0[b], need ? a.block_unit0[b + 1] : a.block_unit0[b], kMergeChunk, l, need,
                        accw);
  ms_finalize<W>(a, st, L, qdeg, ea, d, valid, accw, level, allbits);
}

// First merge stage.  Grid: nmchunk * (BV / THREADS) workgroups.
template <int W, int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_ms_merge_chunk(MsArgs a) {
  constexpr int PER = (1 << VB) / THREADS;
  if (a.ctl->mode != MS_DENSE) return;
  ull allbits[W];
  batch_bits<W>((int)a.ctl->nsrc, allbits);
  ms_merge_chunk<W, VB, 8>(a, a.mchunk[blockIdx.x / PER], (int)(blockIdx.x % PER) * THREADS + (int)threadIdx.x,
                           allbits);
}

// Second merge stage: workgroups take the nitems = nmulti * (BV / THREADS)
// (block, vertex range) items in turn, so that each flushes its per-source
// sums once (thousands of flushes onto the same counters queue on them).
// Grid: at most nitems workgroups.
template <int W, int VB, int THREADS>
__global__ __launch_bounds__(THREADS) void k_ms_merge(MsArgs a, int64_t nitems) {
  constexpr int BV = 1 << VB;
  constexpr int NW = THREADS / kWave;
  constexpr int PER = BV / THREADS;
  __shared__ uint32_t join_ids[NW * kQueue];
  __shared__ uint32_t join_n;
  __shared__ MsStats<W> st;
  if (a.ctl->mode != MS_DENSE) return;
  const uint32_t level = (uint32_t)a.ctl->level + 1;
  ull allbits[W];
  batch_bits<W>((int)a.ctl->nsrc, allbits);
  stats_init<W, THREADS>(st);
  const WgList joins{join_ids, &join_n, (uint32_t)(NW * kQueue)};
  if (threadIdx.x == 0) join_n = 0;
  uint64_t qdeg = 0;
  __syncthreads();
  EdgeArgs ea{};
  ea.off = a.off;
  ea.out_ids = a.list_nxt;
  ea.out_info = &a.ctl->ncount;
  for (int64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    ms_merge_vertex<W, VB, 8>(a, st, joins, qdeg, ea, a.multi[item / PER],
                              (int)(item % PER) * THREADS + (int)threadIdx.x, level, allbits);
    wgl_drain<THREADS>(joins, ea, THREADS);
  }
  wgl_flush<THREADS>(joins, ea, qdeg);
  stats_flush<W, THREADS>(a, st);
}

// ---------------------------------------------------------------------------
// Level bookkeeping.

// After a level: clear the current frontier's words (the buffer becomes the
// next level's output) and its active bits; entries and words [first, ..)
// with stride `step`.
template <int W>
__device__ __forceinline__ void ms_clear(const MsArgs& a, int64_t count, int64_t nwords, int64_t first,
                                         int64_t step) {
  ull* fc = const_cast<ull*>(a.front_cur);
  ull* ac = const_cast<ull*>(a.active_cur);
  for (int64_t i = first; i < count; i += step) {
    const uint32_t v = a.list_cur[i];
#pragma unroll
    for (int w = 0; w < W; ++w) fc[(size_t)v * W + w] = 0;
  }
  for (int64_t i = first; i < nwords; i += step) ac[i] = 0;
}

template <int W>
__global__ void k_ms_clear(MsArgs a, int64_t nwords) {
  if (a.ctl->mode == MS_DONE) return;
  ms_clear<W>(a, a.ctl->fcount, nwords, blockIdx.x * (int64_t)blockDim.x + threadIdx.x,
              (int64_t)gridDim.x * blockDim.x);
}

// Closes a level: per-source reached counts and depths, the next frontier
// becomes current, and the next level's kind is chosen.
__global__ void k_ms_level_end(MsArgs a, ull* reached, int64_t* depth) {
  MsCtl* c = a.ctl;
  if (c->mode == MS_DONE) return;
  const int64_t level = c->level + 1;
  const int nsrc = (int)c->nsrc;
  for (int s = threadIdx.x; s < nsrc; s += blockDim.x) {
    const ull k = a.lvl_new[s];
    if (k) {
      reached[s] += k;
      depth[s] = level;
      a.lvl_new[s] = 0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (c->mode == MS_DENSE) c->dense_levels++;
    else c->sparse_levels++;
    c->level = level;
    c->fcount = c->ncount;
    c->fdeg = c->ndeg;
    c->ncount = 0;
    c->ndeg = 0;
    c->heavy_count = 0;
    c->unit_next = 0;
    c->tag++;
    c->mode = c->fcount == 0 ? MS_DONE : (c->fcount + c->fdeg > c->threshold ? MS_DENSE : MS_SPARSE);
  }
}

// Batch start: source bits, per-source statistics, the first frontier
// (unique source vertices, given in list_cur) and the first level's kind.
__global__ void k_ms_init(MsArgs a, int W, const uint32_t* sources, int64_t nuniq, int64_t threshold,
                          ull* reached, int64_t* depth, uint32_t tag) {
  __shared__ ull degsum;
  if (threadIdx.x == 0) degsum = 0;
  __syncthreads();
  ull* fc = const_cast<ull*>(a.front_cur);
  ull* ac = const_cast<ull*>(a.active_cur);
  for (int s = threadIdx.x; s < a.nsrc; s += blockDim.x) {
    const uint32_t v = sources[s];
    const ull bit = 1ull << (s & 63);
    atomicOr(a.seen + (size_t)v * W + (s >> 6), bit);
    atomicOr(fc + (size_t)v * W + (s >> 6), bit);
    atomicOr(ac + (v >> 6), 1ull << (v & 63));
    reached[s] = 1;
    depth[s] = 0;
    a.edges[s] = a.off[v + 1] - a.off[v];
    a.cs[s] = level_hash(v, 0);
    a.lvl_new[s] = 0;
    if (a.levels) a.levels[(size_t)s * a.n + v] = 0;
    if (a.radii) a.radii[v] = 0;
  }
  for (int64_t i = threadIdx.x; i < nuniq; i += blockDim.x) {
    const uint32_t v = a.list_cur[i];
    atomicAdd(&degsum, (ull)(a.off[v + 1] - a.off[v]));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    MsCtl* c = a.ctl;
    c->level = 0;
    c->fcount = nuniq;
    c->fdeg = (int64_t)degsum;
    c->ncount = c->ndeg = 0;
    c->threshold = threshold;
    c->heavy_count = 0;
    c->unit_next = 0;
    c->dense_levels = c->sparse_levels = 0;
    c->nsrc = a.nsrc;
    c->tag = tag;
    c->mode = c->fcount + c->fdeg > threshold ? MS_DENSE : MS_SPARSE;
  }
}

// ---------------------------------------------------------------------------
// One persistent launch for a whole batch.
//
// A cooperative launch of as many workgroups as can be resident at once
// runs every level of the batch: the same phases as the per-level kernels
// above (sparse: light phase, heavy chunks, vertex pass; dense: units,
// split-block merge; then the clear of the current frontier), separated by
// grid-wide barriers instead of kernel boundaries.  Every workgroup reads
// the level's frontier size after the barrier that closes the level and
// takes the same branch, so all of them meet at every barrier.  A barrier
// wait is bounded by the wall clock: past the limit the waiting workgroup
// sets the abort flag, and every workgroup leaves the level loop at its
// next barrier, so a launch always ends; the host then reports the batch
// as failed.

// Barrier state: workgroup 0 writes `release` (barriers completed) after
// seeing every workgroup's arrival flag; any waiter past the wall-clock
// limit sets `abort`.  Each flag has its own 128-byte line (flags follow
// the control block in the same buffer, cleared before every launch).
struct GridBar {
  unsigned int release;
  unsigned int abort;
  unsigned int pad[30];
};
constexpr int kFlagStride = 32;   // unsigned ints per arrival flag (128 bytes)

// Outputs of one level: the next frontier list (size, out-degree sum, as
// EdgeArgs::out_info), the heavy chunk items queued (sparse) and the unit
// counter (dense).  Three slots rotate: level L writes slot L % 3 while
// workgroup 0 clears slot (L + 1) % 3, which every workgroup last read when
// level L - 1 began.
struct MsSlot {
  int64_t ncount, ndeg;
  int64_t heavy_count;
  unsigned long long unit_next;
  int64_t hook_items;           // BC forward: heavy chunk items of the path-count gathers
  int64_t pad[3];
};

struct MsPersistCtl {
  GridBar bar;
  MsSlot slot[3];
  int64_t pad[8];
};
static_assert(sizeof(MsSlot) == 64 && sizeof(MsPersistCtl) == 384, "persistent control block layout");

struct MsPersistArgs {
  MsPersistCtl* pc;
  unsigned int* flags;          // arrival flags, kFlagStride apart
  const uint32_t* nbr;          // out-neighbours (sparse levels)
  ull* front[2];
  ull* active[2];
  uint32_t* list[2];
  ull* lvl_new[2];              // per-source vertices reached, by level parity
  ull* reached;
  int64_t* depth;
  uint32_t* stamp;
  uint64_t* heavy;              // heavy chunk queue
  int heavy_deg;
  int64_t nunits, nmulti, nmchunk, nwords;
  uint64_t spin_ticks;          // barrier wait limit (wall-clock ticks)
};

// Wall-clock-bounded spin until *p == want; false (and the abort flag set)
// past the limit or once another waiter has set it.
__device__ __forceinline__ bool spin_until(const unsigned int* p, unsigned int want, GridBar* b, uint64_t t0,
                                           uint64_t spin_ticks) {
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
    if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if (wall_clock64() - t0 > spin_ticks) {
      __hip_atomic_store(&b->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
  }
  return true;
}

// Grid-wide barrier.  The workgroup's waves wait for their stores to
// reach the L2, then one thread writes the XCD's L2 back (device-scope
// release fence) and stores the workgroup's arrival flag.  Workgroup 0
// polls every flag with its threads in parallel and stores the release
// word; each workgroup's first thread waits for it and invalidates its
// caches (device-scope acquire fence), so what any workgroup wrote before
// the barrier is seen by all after it.  Fencing in one thread per
// workgroup rather than in every thread keeps a barrier at ~14 us on 256
// workgroups (~170 us with every thread fencing; a counter that all
// workgroups increment costs ~31 us).  Returns false once the barrier has
// been abandoned.
__device__ __forceinline__ bool grid_sync(GridBar* b, unsigned int* flags, unsigned int nwg, unsigned int& gen,
                                          uint64_t spin_ticks, int& s_ok) {
  // Every wave waits until its own global stores have reached the L2: a
  // workgroup-scope release does not wait for them (the workgroup's waves
  // share one CU), and the one L2 write-back below would otherwise miss
  // stores still in flight from other waves.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  const unsigned int g = gen + 1;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(&flags[(size_t)blockIdx.x * kFlagStride], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (blockIdx.x == 0) {
    const uint64_t t0 = wall_clock64();
    for (unsigned int i = threadIdx.x; i < nwg; i += blockDim.x)
      if (!spin_until(&flags[(size_t)i * kFlagStride], g, b, t0, spin_ticks)) break;
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      __hip_atomic_store(&b->release, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (threadIdx.x == 0) {
    (void)spin_until(&b->release, g, b, wall_clock64(), spin_ticks);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = __hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  }
  gen = g;
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return s_ok != 0;
}

// A batch without per-level work beyond the traversal (BFS queries, Radii).
struct MsNoHook {
  static constexpr bool kOn = false;
};

// HOOK = BcPersistHook (one-word batches): at every level, once the next
// frontier is complete and while the current one is intact, the BC forward
// phases (the next level's words and path counts, then the heavy chunks)
// between two more barriers; after the last level, the BC backward levels
// (select, dependencies, heavy chunks) with three barriers each.
template <int W, int VB, int THREADS, class HOOK>
__global__ __launch_bounds__(THREADS) void k_ms_persist(MsArgs a, MsPersistArgs p, HOOK hk) {
  constexpr int BV = 1 << VB;
  constexpr int NW = THREADS / kWave;
  constexpr int PER = BV / THREADS;
  using Op = MsPushOp<W>;
  using Light = LightLds<Op, NW>;
  // Dense levels hold the block's words here; sparse levels, the light
  // phase's tables.
  __shared__ ull lds_main[BV * W];
  __shared__ ull full[BV / 64];
  __shared__ uint8_t fullw[W > 1 ? BV : 1];
  __shared__ uint32_t s_q[NW][kQueue];   // sparse levels' wave queues; dense levels' join list
  __shared__ uint32_t join_n;
  __shared__ MsStats<W> st;
  __shared__ int64_t s_unit;
  __shared__ int s_ok;
  __shared__ double s_w[HOOK::kOn ? 64 : 1];   // BC backward: the workgroup's per-source sums
  __shared__ ull s_nf[HOOK::kOn ? 64 : 1];
  static_assert(sizeof(Light) <= sizeof(ull) * BV * W, "light-phase tables fit in the block's words");
  static_assert(!HOOK::kOn || W == 1, "BC batches use one word of source bits");
  Light& L = *reinterpret_cast<Light*>(lds_main);
  const WgList joins{&s_q[0][0], &join_n, (uint32_t)(NW * kQueue)};
  if (threadIdx.x == 0) join_n = 0;   // ordered by the barriers before the first dense level
  uint64_t qdeg = 0;

  MsCtl* c = a.ctl;
  const int64_t threshold = c->threshold;
  const uint32_t tag0 = c->tag;
  const int nsrc = (int)c->nsrc;
  int64_t fcount = c->fcount, fdeg = c->fdeg;
  ull allbits[W];
  batch_bits<W>(nsrc, allbits);
  const unsigned int nwg = gridDim.x;
  unsigned int gen = 0;
  const int64_t wave_id = (int64_t)blockIdx.x * NW + threadIdx.x / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * NW;
  const int64_t gtid = (int64_t)blockIdx.x * THREADS + threadIdx.x;
  const int64_t gthreads = (int64_t)gridDim.x * THREADS;
  int64_t lev = 0, dense_levels = 0, sparse_levels = 0;
  bool ok = true;
  auto sync = [&]() { return grid_sync(&p.pc->bar, p.flags, nwg, gen, p.spin_ticks, s_ok); };
  for (;; ++lev) {
    const int cur = (int)(lev & 1);
    a.front_cur = p.front[cur];
    a.front_nxt = p.front[cur ^ 1];
    a.active_cur = p.active[cur];
    a.active_nxt = p.active[cur ^ 1];
    a.list_cur = p.list[cur];
    a.list_nxt = p.list[cur ^ 1];
    a.lvl_new = p.lvl_new[cur];
    MsSlot* out = &p.pc->slot[lev % 3];
    if (blockIdx.x == 0) {
      if (lev > 0) {
        // The previous level's counts (flushed before the barrier that closed it).
        ull* prev = p.lvl_new[cur ^ 1];
        for (int s = threadIdx.x; s < nsrc; s += THREADS) {
          const ull k = prev[s];
          if (k) {
            p.reached[s] += k;
            p.depth[s] = lev;
            prev[s] = 0;
          }
        }
      } else {
        for (int s = threadIdx.x; s < nsrc; s += THREADS) p.lvl_new[1][s] = 0;
      }
      if (threadIdx.x < 8) reinterpret_cast<int64_t*>(&p.pc->slot[(lev + 1) % 3])[threadIdx.x] = 0;
    }
    if (fcount == 0) break;
    const bool dense = fcount + fdeg > threshold;
    if (dense) ++dense_levels;
    else ++sparse_levels;
    const uint32_t level = (uint32_t)lev + 1;
    EdgeArgs ea{};
    ea.off = a.off;
    ea.nbr = p.nbr;
    ea.ids = a.list_cur;
    ea.count = fcount;
    ea.out_ids = a.list_nxt;
    ea.out_info = &out->ncount;
    ea.heavy = p.heavy;
    ea.heavy_count = &out->heavy_count;
    ea.heavy_deg = p.heavy_deg;
    WaveQueue q{s_q[threadIdx.x / kWave], 0, 0};
    stats_init<W, THREADS>(st);
    if (dense) {
      ms_dense_units<W, VB, THREADS>(a, p.nunits, &out->unit_next, level, allbits, lds_main, full, fullw, st, joins,
                                     qdeg, ea, s_unit);
      if (!(ok = sync())) break;
      constexpr int MU = W >= 4 ? 2 : 8 / W;
      if (p.nmulti) {
        for (int64_t item = blockIdx.x; item < p.nmchunk * PER; item += gridDim.x)
          ms_merge_chunk<W, VB, MU>(a, a.mchunk[item / PER], (int)(item % PER) * THREADS + (int)threadIdx.x,
                                    allbits);
        if (!(ok = sync())) break;
      }
      for (int64_t item = blockIdx.x; item < p.nmulti * PER; item += gridDim.x) {
        ms_merge_vertex<W, VB, MU>(a, st, joins, qdeg, ea, a.multi[item / PER],
                                   (int)(item % PER) * THREADS + (int)threadIdx.x, level, allbits);
        wgl_drain<THREADS>(joins, ea, THREADS);
      }
      wgl_flush<THREADS>(joins, ea, qdeg);
    } else {
      __syncthreads();
      const Op op{a.front_cur, a.seen, a.next, p.stamp, nullptr, nullptr, tag0 + (uint32_t)lev};
      light_phase<Op, SRC_LIST, HUB_NONE, NW>(op, ea, nullptr, nullptr, q, L, fcount, wave_id, nwaves, false,
                                               (uint64_t)p.heavy_deg);
      if (!(ok = sync())) break;
      heavy_items<Op, HUB_NONE>(op, ea, nullptr, nullptr, q, p.heavy, wave_id, out->heavy_count, nwaves);
      wq_finish(q, ea);
      if (!(ok = sync())) break;
      ms_sparse_fin<W>(a, st, out->ncount, level, gtid, gthreads, allbits);
    }
    if constexpr (HOOK::kOn) {
      if (!(ok = sync())) break;
      const int64_t L1 = lev + 1;
      const int64_t ncount = out->ncount;
      if (L1 >= hk.b.lcap) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && ncount > 0) hk.b.ctl->overflow = 1;
      } else {
        bcf_words(hk.b, L1, a.front_nxt, gtid, gthreads);
        bcf_light(hk.b, ncount, a.list_nxt, a.front_nxt, a.front_cur, a.active_cur, hk.in_off, hk.in_nbr,
                  &out->hook_items, wave_id, nwaves);
      }
      if (!(ok = sync())) break;
      if (L1 < hk.b.lcap)
        bcf_heavy(hk.b, out->hook_items, a.front_nxt, a.front_cur, a.active_cur, hk.in_off, hk.in_nbr, wave_id,
                  nwaves);
      // The heavy chunks read the current frontier's words and bits, which
      // the clear below zeroes.
      if (!(ok = sync())) break;
    }
    ms_clear<W>(a, fcount, p.nwords, gtid, gthreads);
    stats_flush<W, THREADS>(a, st);
    if (!(ok = sync())) break;
    fcount = out->ncount;
    fdeg = out->ndeg;
  }
  if constexpr (HOOK::kOn) {
    // Backward, deepest level (lev - 1) first, unless a source went deeper
    // than the level words hold (the host then runs the batch again).
    if (ok && hk.b.ctl->overflow == 0) {
      const int64_t maxd = lev - 1;
      if (threadIdx.x < 64) {
        s_w[threadIdx.x] = 0.0;
        s_nf[threadIdx.x] = 0;
      }
      double wl = 0.0;
      ull nfl = 0;
      const BcbPSlot* prev = nullptr;
      int64_t j = 0;
      for (int64_t r = maxd; r >= 0; --r, ++j) {
        BcbPSlot* bs = &hk.slots[j % 3];
        if (blockIdx.x == 0 && threadIdx.x < 4) reinterpret_cast<int64_t*>(&hk.slots[(j + 1) % 3])[threadIdx.x] = 0;
        // The heavy vertices of level r + 1 finish before level r reads their D.
        if (prev && r + 1 < maxd) bcb_heavy_fin(hk.b, (uint32_t)(r + 1), prev->hverts, wave_id, nwaves, wl, nfl);
        bcb_select(hk.b, (uint32_t)r, &bs->count, wave_id, nwaves);
        if (!(ok = sync())) break;
        if (r < maxd)
          bcb_light<true>(hk.b, hk.off, hk.nbr, (uint32_t)r, bs->count, &bs->items, &bs->hverts, wave_id, nwaves,
                          wl, nfl);
        else
          bcb_light<false>(hk.b, hk.off, hk.nbr, (uint32_t)r, bs->count, &bs->items, &bs->hverts, wave_id, nwaves,
                           wl, nfl);
        if (!(ok = sync())) break;
        if (r < maxd) bcb_heavy(hk.b, hk.off, hk.nbr, (uint32_t)r, bs->items, wave_id, nwaves);
        if (!(ok = sync())) break;
        prev = bs;
      }
      if (ok) {
        if (prev && maxd > 0) bcb_heavy_fin(hk.b, 0, prev->hverts, wave_id, nwaves, wl, nfl);
        bcb_flush(hk.b, s_w, s_nf, wl, nfl);
      }
    }
  }
  // Every workgroup has read the control block it starts from before the
  // closing values go in.
  (void)sync();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->level = lev;
    c->tag = tag0 + (uint32_t)lev;
    c->fcount = 0;
    c->fdeg = 0;
    c->dense_levels = dense_levels;
    c->sparse_levels = sparse_levels;
    c->mode = MS_DONE;
  }
}

// ---------------------------------------------------------------------------
// Host side.

// Per-context state, grown on demand.  Between batches every buffer except
// seen is all zero (each level clears what it used), so a batch only
// clears its n * W words of seen; `dirty` forces a full clear after a
// failed batch.  The word buffers are sized for the widest batch so far
// (a narrower batch uses their first n * W words).
struct MsState {
  DevBuf seen, front[2], next, stamp, active[2], list[2], part, stats, ctl, srcs, pctl;
  // Resident workgroups of k_ms_persist per word width 1 / 2 / 4 and for
  // BC batches (0: not yet asked).
  int persist_grid[4] = {0, 0, 0, 0};
  int64_t n = 0;
  int wcap = 0;
  uint32_t tag = 0;
  bool dirty = true;
};

void ms_state_free(void* p) { delete static_cast<MsState*>(p); }

int64_t ms_state_bytes(const void* p) {
  const MsState* S = static_cast<const MsState*>(p);
  if (!S) return 0;
  int64_t b = 0;
  for (const DevBuf* d : {&S->seen, &S->front[0], &S->front[1], &S->next, &S->stamp, &S->active[0], &S->active[1],
                          &S->list[0], &S->list[1], &S->part, &S->stats, &S->ctl, &S->srcs, &S->pctl})
    b += (int64_t)d->bytes;
  return b;
}

namespace {

template <class T>
int zero(rg_ctx* ctx, DevBuf& b) {
  if (b.bytes) RG_TRY(fill<T>(ctx, b.as<T>(), (int64_t)(b.bytes / sizeof(T)), T(0)));
  return RG_OK;
}

int ms_prepare(rg_ctx* ctx, MsState& S, int64_t n, int W, int64_t nunits) {
  if (S.n != n || W > S.wcap) {
    const int wcap = std::max(W, S.n == n ? S.wcap : 0);
    const size_t words = (size_t)n * wcap * sizeof(ull);
    const size_t aw = sizeof(ull) * (size_t)(n / 64 + 1);
    for (DevBuf* b : {&S.seen, &S.front[0], &S.front[1], &S.next, &S.stamp, &S.active[0], &S.active[1],
                      &S.list[0], &S.list[1]})
      b->release();
    RG_TRY(S.seen.ensure(words));
    RG_TRY(S.front[0].ensure(words));
    RG_TRY(S.front[1].ensure(words));
    RG_TRY(S.next.ensure(words));
    RG_TRY(S.stamp.ensure(sizeof(uint32_t) * n));
    RG_TRY(S.active[0].ensure(aw));
    RG_TRY(S.active[1].ensure(aw));
    RG_TRY(S.list[0].ensure(sizeof(uint32_t) * n));
    RG_TRY(S.list[1].ensure(sizeof(uint32_t) * n));
    S.n = n;
    S.wcap = wcap;
    S.dirty = true;
  }
  const size_t pb = ((size_t)std::max<int64_t>(nunits, 1) << kMsVB) * W * sizeof(ull);
  if (S.part.bytes < pb) {
    S.part.release();
    RG_TRY(S.part.ensure(pb));
  }
  // lvl_new, edges, cs, reached, depth, and lvl_new of odd levels (persistent launch)
  RG_TRY(S.stats.ensure(sizeof(ull) * 6 * kMsMaxSources));
  RG_TRY(S.ctl.ensure(sizeof(MsCtl)));
  RG_TRY(S.pctl.ensure(sizeof(MsPersistCtl) + sizeof(unsigned int) * kFlagStride * 4 * 1024));
  RG_TRY(S.srcs.ensure(sizeof(uint32_t) * kMsMaxSources));
  if (S.dirty || S.tag > 0x7fffffffu) {
    for (DevBuf* b : {&S.front[0], &S.front[1], &S.next, &S.active[0], &S.active[1]}) RG_TRY(zero<ull>(ctx, *b));
    RG_TRY(zero<uint32_t>(ctx, S.stamp));
    S.tag = 0;
    S.dirty = false;
  }
  return RG_OK;
}

// One persistent batch in the process at a time, whatever its width or
// hook: two cooperative grids from different contexts, each partly
// resident, would wait on each other's workgroups until the barrier limit.
std::mutex g_persist_mu;

// Runs the batch that k_ms_init set up as one cooperative launch of
// k_ms_persist: as many workgroups as the device holds at once (the launch
// fails rather than runs if they cannot all be resident).
template <int W, class HOOK>
int ms_persist(rg_graph* g, MsState& S, BlockedCSC* B, const Engine& E, const MsArgs& a, ull* reached,
               int64_t* depth, ull* lvl_new_odd, int64_t* levels_run, const HOOK& hk) {
  rg_ctx* ctx = g->ctx;
  constexpr int T = ms_threads<W>();
  auto kern = k_ms_persist<W, kMsVB, T, HOOK>;
  int& grid = S.persist_grid[HOOK::kOn ? 3 : W == 1 ? 0 : W == 2 ? 1 : 2];
  if (grid == 0) {
    int coop = 0;
    RG_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, ctx->device));
    if (!coop) return set_error(RG_EUNSUPPORTED, "persistent batch: no cooperative launch on this device");
    int per_cu = 0;
    RG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, T, 0));
    if (per_cu < 1) return set_error(RG_EUNSUPPORTED, "persistent batch: the kernel does not fit a CU");
    grid = per_cu * ctx->num_cus;
  }
  int rate_khz = 0;
  RG_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (grid > 4 * 1024) return set_error(RG_EUNSUPPORTED, "persistent batch: more than 4096 resident workgroups");
  MsPersistArgs p{};
  p.pc = S.pctl.as<MsPersistCtl>();
  p.flags = reinterpret_cast<unsigned int*>(p.pc + 1);
  p.nbr = g->nbr;
  for (int i = 0; i < 2; ++i) {
    p.front[i] = S.front[i].as<ull>();
    p.active[i] = S.active[i].as<ull>();
    p.list[i] = S.list[i].as<uint32_t>();
  }
  p.lvl_new[0] = a.lvl_new;
  p.lvl_new[1] = lvl_new_odd;
  p.reached = reached;
  p.depth = depth;
  p.stamp = S.stamp.as<uint32_t>();
  p.heavy = E.heavy;
  p.heavy_deg = E.heavy_deg;
  p.nunits = B->nunits;
  p.nmulti = B->nmulti;
  p.nmchunk = B->nmchunk;
  p.nwords = g->n / 64 + 1;
  p.spin_ticks = (uint64_t)std::max(rate_khz, 1) * 1000ull * 10ull;   // a barrier wait of 10 s
  RG_TRY(fill<ull>(ctx, reinterpret_cast<ull*>(p.pc),
                   (int64_t)((sizeof(MsPersistCtl) + sizeof(unsigned int) * kFlagStride * grid) / sizeof(ull)), 0ull));
  MsArgs ka = a;
  HOOK kh = hk;
  void* args[] = {&ka, &p, &kh};
  // The lock is held until the launch has ended (the reads below wait for it).
  std::lock_guard<std::mutex> lock(g_persist_mu);
  RG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(grid), dim3(T), args, 0, ctx->stream));
  int64_t bar[1] = {0};   // GridBar: release | abort
  int64_t lev = 0;
  RG_TRY(ctx_read_count(ctx, reinterpret_cast<const int64_t*>(p.pc), bar, 1));
  RG_TRY(ctx_read_count(ctx, &a.ctl->level, &lev, 1));
  if ((uint32_t)((uint64_t)bar[0] >> 32) != 0)
    return set_error(RG_EDEVICE, "persistent batch: a grid barrier wait ran past its limit");
  S.tag += (uint32_t)lev + 1;
  S.dirty = false;
  if (levels_run) *levels_run = lev;
  return RG_OK;
}

template <int W, class HOOK = MsNoHook>
int ms_run(rg_graph* g, MsState& S, BlockedCSC* B, const Engine& E, const uint32_t* sources, int nsrc,
           int64_t threshold, int32_t* levels_dev, int32_t* radii_dev, int64_t* levels_run, MsLevelHook* hook,
           bool persist, const HOOK& dh = HOOK{}) {
  rg_ctx* ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = g->n;
  const int64_t nwords = n / 64 + 1;
  RG_TRY(fill<ull>(ctx, S.seen.as<ull>(), n * W, 0ull));
  // Unique source vertices form the first frontier list.
  std::vector<uint32_t> uniq(sources, sources + nsrc);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  RG_HIP(hipMemcpyAsync(S.srcs.p, sources, sizeof(uint32_t) * nsrc, hipMemcpyHostToDevice, st));
  RG_HIP(hipMemcpyAsync(S.list[0].p, uniq.data(), sizeof(uint32_t) * uniq.size(), hipMemcpyHostToDevice, st));
  ull* stats = S.stats.as<ull>();
  MsArgs a{};
  a.ctl = S.ctl.as<MsCtl>();
  a.n = n;
  a.nsrc = nsrc;
  a.off = g->off;
  a.src = B->src;
  a.vl = B->vl;
  a.unit_lo = B->unit_lo;
  a.unit_block = B->unit_block;
  a.block_unit0 = B->block_unit0;
  a.multi = B->multi;
  a.mchunk = B->mchunk;
  a.seen = S.seen.as<ull>();
  a.next = S.next.as<ull>();
  a.part = S.part.as<ull>();
  a.lvl_new = stats;
  a.edges = stats + kMsMaxSources;
  a.cs = stats + 2 * kMsMaxSources;
  ull* reached = stats + 3 * kMsMaxSources;
  int64_t* depth = reinterpret_cast<int64_t*>(stats + 4 * kMsMaxSources);
  a.levels = levels_dev;
  a.radii = radii_dev;
  // RG_BATCH_WIDE_STATS forces the statistics of graphs with 2^32 arcs or more.
  a.wide = g->m >= (int64_t(1) << 32) || std::getenv("RG_BATCH_WIDE_STATS") ? 1u : 0u;
  int cur = 0;
  auto bind = [&](int c) {
    a.front_cur = S.front[c].as<ull>();
    a.front_nxt = S.front[c ^ 1].as<ull>();
    a.active_cur = S.active[c].as<ull>();
    a.active_nxt = S.active[c ^ 1].as<ull>();
    a.list_cur = S.list[c].as<uint32_t>();
    a.list_nxt = S.list[c ^ 1].as<uint32_t>();
  };
  bind(cur);
  S.dirty = true;   // until the batch completes
  hipLaunchKernelGGL(k_ms_init, 1, 256, 0, st, a, W, S.srcs.as<uint32_t>(), (int64_t)uniq.size(), threshold,
                     reached, depth, S.tag + 1u);
  RG_HIP(hipGetLastError());
  if (persist && !hook)
    return ms_persist<W, HOOK>(g, S, B, E, a, reached, depth, stats + 5 * kMsMaxSources, levels_run, dh);

  EdgeArgs ea{};
  ea.off = g->off;
  ea.nbr = g->nbr;
  // The list size is read on the device; the bound (n) only sizes the grid,
  // the full persistent one: sparse levels are bound by atomics in flight.
  ea.count = n;
  ea.count_dev = &a.ctl->fcount;
  ea.out_info = &a.ctl->ncount;
  ea.heavy = E.heavy;
  ea.heavy_count = &a.ctl->heavy_count;
  ea.heavy_deg = E.heavy_deg;
  const int gfin = std::max(1, ctx->num_cus * 4);
  // Dense level: a persistent grid taking units from a counter (the LDS
  // copy limits how many run per CU: 4 for W = 1, 2 for W = 2, 1 for W = 4).
  const int gdense = (int)std::min<int64_t>(B->nunits, (int64_t)ctx->num_cus * 4);
  const int gclr = grid_for(ctx, std::max<int64_t>(nwords, 1));
  constexpr int kMsThreads = ms_threads<W>();
  constexpr int PER = (1 << kMsVB) / kMsThreads;
  // One chunk of kMsLevelsPerSync levels (an even number: the chunk starts
  // and ends on the same buffer parity, and its kernels read the level, the
  // stamp tag and the source count from MsCtl, so every chunk of a batch,
  // and of every batch of this word width on these buffers, is the same
  // launch sequence).
  auto chunk = [&]() -> int {
    for (int k = 0; k < kMsLevelsPerSync; ++k) {
      MsPushOp<W> op{a.front_cur, a.seen, a.next, S.stamp.as<uint32_t>(), &a.ctl->tag, &a.ctl->mode, 0u};
      ea.ids = a.list_cur;
      ea.out_ids = a.list_nxt;
      RG_TRY((launch_hub<MsPushOp<W>, SRC_LIST>(E, op, ea)));
      hipLaunchKernelGGL(k_ms_sparse_fin<W>, gfin, 256, 0, st, a);
      hipLaunchKernelGGL((k_ms_dense<W, kMsVB, kMsThreads>), gdense, kMsThreads, 0, st, a, B->nunits);
      if (B->nmulti) {
        hipLaunchKernelGGL((k_ms_merge_chunk<W, kMsVB, kMsThreads>), (int)(B->nmchunk * PER), kMsThreads, 0, st, a);
        hipLaunchKernelGGL((k_ms_merge<W, kMsVB, kMsThreads>),
                           (int)std::min<int64_t>(B->nmulti * PER, (int64_t)ctx->num_cus * 2), kMsThreads, 0, st, a,
                           B->nmulti * PER);
      }
      if (hook) {
        const MsLevelView view{&a.ctl->mode, &a.ctl->level, &a.ctl->ncount, a.list_nxt,
                               a.front_cur,  a.front_nxt,   a.active_cur};
        RG_TRY(hook->level(view));
      }
      hipLaunchKernelGGL(k_ms_clear<W>, gclr, 256, 0, st, a, nwords);
      hipLaunchKernelGGL(k_ms_level_end, 1, 256, 0, st, a, reached, depth);
      RG_HIP(hipGetLastError());
      cur ^= 1;
      bind(cur);
    }
    return RG_OK;
  };
  // Without a level hook and with per-launch timing off, the chunk is
  // captured once as a HIP graph (per word width, while the buffers, layout
  // and engine stay the same) and replayed.
  const bool captured = !hook && ctx->graphs && !ctx->timing;
  GraphCache& gc = ctx->ms_graph[W == 1 ? 0 : W == 2 ? 1 : 2];
  if (captured) {
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)g->off, (uint64_t)(uintptr_t)g->nbr, (uint64_t)(uintptr_t)B->src,
        (uint64_t)(uintptr_t)B->vl, (uint64_t)(uintptr_t)B->unit_lo, (uint64_t)(uintptr_t)B->unit_block,
        (uint64_t)(uintptr_t)B->block_unit0, (uint64_t)(uintptr_t)B->multi, (uint64_t)(uintptr_t)B->mchunk,
        (uint64_t)B->nmchunk, (uint64_t)a.wide, (uint64_t)B->nunits,
        (uint64_t)B->nmulti, (uint64_t)n, (uint64_t)(uintptr_t)a.ctl, (uint64_t)(uintptr_t)a.seen,
        (uint64_t)(uintptr_t)a.next, (uint64_t)(uintptr_t)a.part, (uint64_t)(uintptr_t)stats,
        (uint64_t)(uintptr_t)S.front[0].p, (uint64_t)(uintptr_t)S.front[1].p, (uint64_t)(uintptr_t)S.active[0].p,
        (uint64_t)(uintptr_t)S.active[1].p, (uint64_t)(uintptr_t)S.list[0].p, (uint64_t)(uintptr_t)S.list[1].p,
        (uint64_t)(uintptr_t)S.stamp.p, (uint64_t)(uintptr_t)E.heavy, (uint64_t)E.heavy_deg, (uint64_t)E.block,
        (uint64_t)(uintptr_t)levels_dev, (uint64_t)(uintptr_t)radii_dev};
    if (gc.key != key || !gc.exec) {
      gc.release();
      RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = chunk();
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &graph);
      if (rc != RG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      RG_HIP(ec);
      gc.graph = graph;
      RG_HIP(hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0));
      gc.key = key;
    }
  }
  int64_t steps = 0;
  for (;;) {
    if (captured) RG_HIP(hipGraphLaunch(gc.exec, st));
    else RG_TRY(chunk());
    steps += kMsLevelsPerSync;
    int64_t mode = 0;
    RG_TRY(ctx_read_count(ctx, &a.ctl->mode, &mode, 1));
    if (mode == MS_DONE) break;
    if (steps > n + 8) return set_error(RG_EDEVICE, "multi-source BFS: no convergence");
  }
  S.tag += (uint32_t)steps + 1;
  S.dirty = false;
  if (levels_run) *levels_run = steps;
  return RG_OK;
}

}  // namespace

int bfs_batch_prepare(rg_graph* g) {
  BlockedCSC* B = nullptr;
  return query_layout(g, &B);
}

int bfs_batch_reserve(rg_graph* g, int nsrc) {
  rg_ctx* ctx = g->ctx;
  BlockedCSC* B = nullptr;
  RG_TRY(query_layout(g, &B));
  MsState* S = static_cast<MsState*>(ctx->ms);
  if (!S) {
    S = new MsState;
    ctx->ms = S;
  }
  const int W = nsrc <= 64 ? 1 : nsrc <= 128 ? 2 : 4;
  return ms_prepare(ctx, *S, g->n, W, B->nunits);
}

int64_t bfs_batch_bytes(int64_t n, int64_t m, int nsrc) {
  const int64_t W = nsrc <= 64 ? 1 : nsrc <= 128 ? 2 : 4;
  const int64_t units = (n >> kMsVB) + m / kUnitEdges + 2;
  return 4 * n * W * 8 + (units << kMsVB) * W * 8 + n * 12 + n / 4 + 4096;
}

// Launch form of a batch (rg_opts::batch_launch): 1 = per-level kernels,
// 2 = one persistent launch; 0 = the library's choice: the persistent
// launch from 2^27 arcs.  Measured on MI355X (device time per batch of 64 /
// 256 sources, per-level vs persistent): symmetric R-MAT of 2^18 vertices
// 1.51 / 2.68 vs 1.83 / 2.75 ms, 2^22: 2.64 / 5.55 vs 2.89 / 5.53 ms, 2^26:
// 27.4 / 81.5 vs 25.4 / 79.0 ms -- a grid barrier (~14 us) costs more than
// a kernel boundary inside a HIP graph (~1.6 us), which small graphs feel,
// while on large ones the single launch saves the host checks and its
// dense phase runs without the per-level kernel's register spills.
static bool batch_persistent(const rg_graph* g, const rg_opts* o) {
  const int mode = o ? o->batch_launch : 0;
  if (mode == 1) return false;
  if (mode == 2) return true;
  return g->m >= ((int64_t)1 << 27);
}

namespace {

// Runs one batch; per-source summaries into out (nullable), level arrays or
// radii into the device buffers given (nullable).
int batch_run(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, int32_t* levels_dev,
              int32_t* radii_dev, rg_bfs_summary* out, double* ms_out, MsLevelHook* hook = nullptr,
              const BcPersistHook* bc = nullptr) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  const int W = nsrc <= 64 ? 1 : nsrc <= 128 ? 2 : 4;
  BlockedCSC* B = nullptr;
  RG_TRY(query_layout(g, &B));
  Engine E;
  rg_opts po;
  rg_opts_default(&po);
  if (o) {
    po.heavy_degree = o->heavy_degree;
    po.wg_size = o->wg_size;
  }
  po.variant = RG_PUSH;
  RG_TRY(make_engine(g, &po, sizeof(uint32_t), E, 4));
  const int64_t threshold = (o && o->threshold >= 0) ? o->threshold : g->m / 20;
  MsState* S = static_cast<MsState*>(ctx->ms);
  if (!S) {
    S = new MsState;
    ctx->ms = S;
  }
  RG_TRY(ms_prepare(ctx, *S, n, W, B->nunits));
  EventTimer tm(ctx->stream);
  int64_t steps = 0;
  const bool persist = !hook && (bc || batch_persistent(g, o));
  int rc = bc       ? ms_run<1, BcPersistHook>(g, *S, B, E, sources, nsrc, threshold, levels_dev, radii_dev, &steps,
                                               nullptr, true, *bc)
           : W == 1 ? ms_run<1>(g, *S, B, E, sources, nsrc, threshold, levels_dev, radii_dev, &steps, hook, persist)
           : W == 2 ? ms_run<2>(g, *S, B, E, sources, nsrc, threshold, levels_dev, radii_dev, &steps, hook, persist)
                    : ms_run<4>(g, *S, B, E, sources, nsrc, threshold, levels_dev, radii_dev, &steps, hook, persist);
  if (rc != RG_OK) return rc;
  double ms = tm.stop_ms();
  if (out) {
    std::vector<ull> h(5 * kMsMaxSources);
    RG_TRY(ctx_copy(ctx, h.data(), S->stats.p, sizeof(ull) * h.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < nsrc; ++s) {
      out[s].iters = (int64_t)h[4 * kMsMaxSources + s] + 1;
      out[s].reached = (int64_t)h[3 * kMsMaxSources + s];
      out[s].edges = (int64_t)h[kMsMaxSources + s];
      out[s].checksum = h[2 * kMsMaxSources + s];
    }
  }
  if (ms_out) *ms_out = ms;
  return RG_OK;
}

}  // namespace

int bfs_batch(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, int32_t* levels_out,
              rg_bfs_summary* out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (nsrc < 1 || nsrc > kMsMaxSources)
    return set_error(RG_EINVAL, "rg_bfs_batch: between 1 and 256 sources per batch");
  if (!sources || !out) return set_error(RG_EINVAL, "rg_bfs_batch: null argument");
  for (int i = 0; i < nsrc; ++i)
    if ((int64_t)sources[i] >= n) return set_error(RG_EINVAL, "rg_bfs_batch: source out of range");
  if (n >= (int64_t)1 << 32) return set_error(RG_EUNSUPPORTED, "rg_bfs_batch: n >= 2^32");
  DevBuf lv;
  int32_t* levels_dev = nullptr;
  if (levels_out) {
    RG_TRY(lv.ensure(sizeof(int32_t) * (size_t)nsrc * n));
    levels_dev = lv.as<int32_t>();
    RG_TRY(fill<int32_t>(ctx, levels_dev, (int64_t)nsrc * n, -1));
  }
  RG_TRY(batch_run(g, sources, nsrc, o, levels_dev, nullptr, out, ms_out));
  if (levels_out)
    RG_TRY(ctx_copy(ctx, levels_out, levels_dev, sizeof(int32_t) * (size_t)nsrc * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

int bfs_batch_hooked(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, MsLevelHook* hook,
                     rg_bfs_summary* out) {
  if (nsrc < 1 || nsrc > 64) return set_error(RG_EINVAL, "hooked batch: between 1 and 64 sources");
  return batch_run(g, sources, nsrc, o, nullptr, nullptr, out, nullptr, hook);
}

int bfs_batch_bc_persist(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, const BcPersistHook& hk,
                         rg_bfs_summary* fw) {
  if (nsrc < 1 || nsrc > 64) return set_error(RG_EINVAL, "BC batch: between 1 and 64 sources");
  return batch_run(g, sources, nsrc, o, nullptr, nullptr, fw, nullptr, nullptr, &hk);
}

// Radii (Radii.C:98-157) as one bit-parallel batch: radii[v] = the last
// level at which v gained a source bit (max distance from the starts), -1
// if unreached; rounds = levels with a nonempty frontier.  Only the set of
// start vertices matters (a repeated start in the reference overwrites its
// own word, Radii.C:124-126), so starts are deduplicated.
int radii_batch(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o, int32_t* radii_out,
                int64_t* rounds_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (nstarts < 0 || nstarts > 64) return set_error(RG_EINVAL, "rg_radii: 0..64 starts");
  for (int i = 0; i < nstarts; ++i)
    if ((int64_t)starts[i] >= n) return set_error(RG_EINVAL, "rg_radii: start out of range");
  std::vector<uint32_t> uniq(starts, starts + nstarts);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  DevBuf& b_rad = ctx->scratch[0];
  RG_TRY(b_rad.ensure(sizeof(int32_t) * n));
  int32_t* radii = b_rad.as<int32_t>();
  RG_TRY(fill<int32_t>(ctx, radii, n, -1));
  int64_t rounds = 0;
  double ms = 0.0;
  if (!uniq.empty()) {
    std::vector<rg_bfs_summary> sum(uniq.size());
    RG_TRY(batch_run(g, uniq.data(), (int)uniq.size(), o, nullptr, radii, sum.data(), &ms));
    for (const auto& x : sum) rounds = std::max(rounds, x.iters);
  }
  if (rounds_out) *rounds_out = rounds;
  if (ms_out) *ms_out = ms;
  if (radii_out) RG_TRY(ctx_copy(ctx, radii_out, radii, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

}  // namespace rg

extern "C" {

int rg_bfs_batch(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, int32_t* levels_out,
                 rg_bfs_summary* out, double* ms_out) {
  if (!g) return rg::set_error(RG_EINVAL, "rg_bfs_batch: null graph");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::bfs_batch(g, sources, nsources, o, levels_out, out, ms_out);
}

uint64_t rg_level_hash(uint32_t v, uint32_t level) { return rg::level_hash(v, level); }

}  // extern "C"
