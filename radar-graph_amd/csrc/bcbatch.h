// Device side of the batched betweenness centrality (one batch of up to 64
// single-source queries, source s owning bit s of a 64-bit word): the
// phases of a forward level (the next level's words and path counts) and
// of a backward level (the level's vertices, their dependencies), as
// functions over a range of waves.  bcbatch.hip launches each phase as a
// kernel; the persistent batch kernel (msbfs.hip) runs them between grid
// barriers.  The algorithm is described in bcbatch.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device.h"
#include "engine.h"
#include "queries.h"

namespace rg {

constexpr int kBcbMaxSources = 64;    // sources per batch (one word of source bits)
constexpr int kBcbMaxBatch = 256;     // sources per call, in batches of 64
// A batch of fewer sources costs more than running them one by one: the
// bit-parallel forward and the per-level gathers scan every arc whatever
// the width (symmetric R-MAT of 2^25 vertices, RADAR engine: a batch of 1
// / 37 / 64 takes 47 / 212 / 256 ms against 8.8 ms per single-source run;
// 2^26: 1 / 8 / 64 take 99 / 273 / 489 ms against 22 ms).
constexpr int kBcbSingleBelow = 12;
constexpr int kBcbMaxLevels = 64;      // levels 0..63 kept; deeper sources take single-source runs
constexpr int kBcbLevelsInit = 32;     // level words allocated at first (grown on demand)
constexpr uint64_t kBcbHeavy = 4096;   // degree above which edges are split into chunks
constexpr uint64_t kBcbChunk = 4096;
constexpr int kBcbQ = 4;               // hits whose rows are loaded together
// Vertices per wave in the light kernels.  One: the gathers are bound by
// the bytes of the rows they read, and more vertices per wave (more load
// chains in flight, more registers) measured slower at 2 and 4.
constexpr int kBcbV = 1;
constexpr int kBcbThreads = 256;   // per-level kernels' workgroup
constexpr int kBcbWaves = kBcbThreads / kWave;

typedef unsigned long long ull;

struct BcbCtl {
  int64_t items;     // heavy chunk items queued at this level
  int64_t hverts;    // heavy vertices of this backward level
  int64_t count;     // vertices of this backward level
  int64_t overflow;  // a source reached a level beyond the level words allocated
};

struct BcbArgs {
  BcbCtl* ctl;
  int64_t n;
  int nsrc;
  int lcap;              // levels with words allocated
  double* sig;           // [n][64] path counts, then D
  ull* lw;               // [lcap][n] per level: sources for which the vertex is at that level
  ull* items;            // heavy chunk queue: (vertex or heavy slot) << 32 | chunk
  uint32_t* hv;          // backward: heavy vertices of the level
  double* acc_h;         // backward: [heavy slot][64] sums
  uint32_t* list;        // backward: vertices of the level
  double* wsum;          // [64] weighted sums of the dependencies
  ull* nonfin;           // [64] non-finite dependencies among the reached vertices
  double* deps;          // optional [nsrc][n]
};

__device__ __forceinline__ uint32_t bcb_rl32(uint32_t x, int j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}
__device__ __forceinline__ ull bcb_rl64(ull x, int j) {
  return ((ull)bcb_rl32((uint32_t)(x >> 32), j) << 32) | bcb_rl32((uint32_t)x, j);
}

// For V edge ranges at once (V vertices, or V parts of one): acc[t] +=
// rows[u][lane] over the edges [e0[t], e1[t]) of the neighbour list nbr
// whose neighbour u has bit `lane` of mask[t] in its word words[u].
// FILTER: test the bitmap `active` (words[u] != 0) before loading the word.
// Per step a lane takes one edge of each range (V independent load chains
// in flight); the hits of each range are then taken four at a time with a
// lane per source, each reading its own word of the neighbour's 64-word
// row.  Called by whole waves with wave-uniform ranges and masks.
template <int V, bool FILTER>
__device__ __forceinline__ void gather_rows(const uint32_t* nbr, const uint64_t* e0, const uint64_t* e1,
                                            const ull* mask, const ull* active, const ull* words,
                                            const double* rows, int lane, double* acc) {
  uint64_t len = 0;
#pragma unroll
  for (int t = 0; t < V; ++t) len = std::max<uint64_t>(len, e1[t] - e0[t]);
  for (uint64_t o = 0; o < len; o += kWave) {
    uint32_t u[V];
    bool in[V];
#pragma unroll
    for (int t = 0; t < V; ++t) {
      const uint64_t k = e0[t] + o + lane;
      in[t] = k < e1[t];
      u[t] = in[t] ? nbr[k] : 0u;
    }
    if (FILTER) {
#pragma unroll
      for (int t = 0; t < V; ++t) in[t] = in[t] && ((active[u[t] >> 6] >> (u[t] & 63)) & 1);
    }
    ull f[V];
#pragma unroll
    for (int t = 0; t < V; ++t) f[t] = in[t] ? (words[u[t]] & mask[t]) : 0ull;
#pragma unroll
    for (int t = 0; t < V; ++t) {
      ull hit = __ballot(f[t] != 0);
      while (hit) {
        uint32_t uu[kBcbQ];
        bool on[kBcbQ];
#pragma unroll
        for (int q = 0; q < kBcbQ; ++q) {
          const bool have = hit != 0;
          const int j = have ? __ffsll((long long)hit) - 1 : 0;
          hit &= hit - 1;
          const ull fj = bcb_rl64(f[t], j);
          uu[q] = bcb_rl32(u[t], j);
          on[q] = have && ((fj >> lane) & 1);
        }
        double x[kBcbQ];
#pragma unroll
        for (int q = 0; q < kBcbQ; ++q) x[q] = on[q] ? rows[(size_t)uu[q] * 64 + lane] : 0.0;
#pragma unroll
        for (int q = 0; q < kBcbQ; ++q) acc[t] += x[q];
      }
    }
  }
}

// A heavy chunk [lo, hi) as four ranges gathered together.
template <bool FILTER>
__device__ __forceinline__ double gather_chunk(const uint32_t* nbr, uint64_t lo, uint64_t hi, ull mask,
                                               const ull* active, const ull* words, const double* rows, int lane) {
  constexpr uint64_t kPart = kBcbChunk / 4;
  uint64_t e0[4], e1[4];
  ull mk[4];
  double acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    e0[t] = std::min<uint64_t>(lo + t * kPart, hi);
    e1[t] = std::min<uint64_t>(lo + (t + 1) * kPart, hi);
    mk[t] = mask;
    acc[t] = 0.0;
  }
  gather_rows<4, FILTER>(nbr, e0, e1, mk, active, words, rows, lane, acc);
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// ---------------------------------------------------------------------------
// Forward: level L -> L + 1 (L1 = L + 1 < lcap).

// lw[L1] = the next frontier's words (every vertex written); entries
// [first, n) with stride `step`.
__device__ __forceinline__ void bcf_words(const BcbArgs& b, int64_t L1, const ull* front_nxt, int64_t first,
                                          int64_t step) {
  ull* dst = b.lw + L1 * b.n;
  for (int64_t i = first; i < b.n; i += step) dst[i] = front_nxt[i];
}

// kBcbV vertices of the next frontier (list_nxt[0..count)) per wave: their
// path counts.  A vertex of in-degree above kBcbHeavy has its entries
// zeroed here and its chunks queued (counter *items) for bcf_heavy.
__device__ __forceinline__ void bcf_light(const BcbArgs& b, int64_t count, const uint32_t* list_nxt,
                                          const ull* front_nxt, const ull* front_cur, const ull* active_cur,
                                          const uint64_t* in_off, const uint32_t* in_nbr, int64_t* items,
                                          int64_t wave, int64_t nwaves) {
  const int lane = (int)__lane_id();
  for (int64_t i0 = wave * kBcbV; i0 < count; i0 += nwaves * kBcbV) {
    uint32_t d[kBcbV];
    ull nn[kBcbV];
    uint64_t e0[kBcbV], e1[kBcbV];
    double acc[kBcbV];
#pragma unroll
    for (int t = 0; t < kBcbV; ++t) {
      const bool valid = i0 + t < count;
      d[t] = valid ? list_nxt[i0 + t] : 0u;
      nn[t] = valid ? front_nxt[d[t]] : 0ull;
      e0[t] = valid ? in_off[d[t]] : 0;
      e1[t] = valid ? in_off[d[t] + 1] : 0;
      acc[t] = 0.0;
      if (e1[t] - e0[t] > kBcbHeavy) {
        if ((nn[t] >> lane) & 1) b.sig[(size_t)d[t] * 64 + lane] = 0.0;
        const uint64_t nch = (e1[t] - e0[t] + kBcbChunk - 1) / kBcbChunk;
        ull base = 0;
        if (lane == 0) base = atomicAdd((ull*)items, (ull)nch);
        base = bcb_rl64(base, 0);
        for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = ((ull)d[t] << 32) | c;
        nn[t] = 0;          // nothing more to write for this vertex here
        e1[t] = e0[t];
      }
    }
    gather_rows<kBcbV, true>(in_nbr, e0, e1, nn, active_cur, front_cur, b.sig, lane, acc);
#pragma unroll
    for (int t = 0; t < kBcbV; ++t)
      if ((nn[t] >> lane) & 1) b.sig[(size_t)d[t] * 64 + lane] = acc[t];
  }
}

// The chunk items [0, count) that bcf_light queued: sums met by FP64 atomics.
__device__ __forceinline__ void bcf_heavy(const BcbArgs& b, int64_t count, const ull* front_nxt,
                                          const ull* front_cur, const ull* active_cur, const uint64_t* in_off,
                                          const uint32_t* in_nbr, int64_t wave, int64_t nwaves) {
  const int lane = (int)__lane_id();
  for (int64_t i = wave; i < count; i += nwaves) {
    const ull it = b.items[i];
    const uint32_t d = (uint32_t)(it >> 32);
    const uint64_t c = it & 0xffffffffull;
    const ull nn = front_nxt[d];
    const uint64_t lo = in_off[d] + c * kBcbChunk;
    const uint64_t hi = std::min<uint64_t>(in_off[d + 1], lo + kBcbChunk);
    const double s = gather_chunk<true>(in_nbr, lo, hi, nn, active_cur, front_cur, b.sig, lane);
    if (((nn >> lane) & 1) && s != 0.0) unsafeAtomicAdd(b.sig + (size_t)d * 64 + lane, s);
  }
}

// ---------------------------------------------------------------------------
// Backward: level r, from the deepest down to 0.

// Vertices at level r from some source into b.list (size *count).  Each
// wave scans a contiguous range of 64-vertex words twice: counting, one
// reservation in the list, writing.
__device__ __forceinline__ void bcb_select(const BcbArgs& b, uint32_t r, int64_t* count, int64_t wave,
                                           int64_t nwaves) {
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  const int64_t nwords = (b.n + kWave - 1) / kWave;
  const int64_t per = (nwords + nwaves - 1) / nwaves;
  const int64_t w0 = wave * per, w1 = std::min<int64_t>(w0 + per, nwords);
  ull total = 0;
  for (int64_t w = w0; w < w1; ++w) {
    const int64_t v = w * kWave + lane;
    total += __popcll(__ballot(v < b.n && lr[v] != 0));
  }
  if (!total) return;
  ull pos = 0;
  if (lane == 0) pos = atomicAdd((ull*)count, total);
  pos = bcb_rl64(pos, 0);
  for (int64_t w = w0; w < w1; ++w) {
    const int64_t v = w * kWave + lane;
    const bool at = v < b.n && lr[v] != 0;
    const ull word = __ballot(at);
    if (at) b.list[pos + __popcll(word & ((1ull << lane) - 1))] = (uint32_t)v;
    pos += __popcll(word);
  }
}

// D and the dependency of (v, lane) from the sum over its children.
__device__ __forceinline__ void bwd_finish(const BcbArgs& b, uint32_t v, int lane, double acc, double& wl,
                                           ull& nfl) {
  const size_t row = (size_t)v * 64 + lane;
  const double inv = 1.0 / b.sig[row];
  const double D = acc + inv;
  b.sig[row] = D;
  const double dep = (D - inv) / inv;
  if (isfinite(dep)) wl += dep * summary_weight(v);
  else nfl++;
  if (b.deps) b.deps[(size_t)lane * b.n + v] = dep;
}

// Per-lane sums of a workgroup into the per-source totals (s_w, s_nf: the
// workgroup's 64 LDS entries each, zeroed before the sums began).
__device__ __forceinline__ void bcb_flush(const BcbArgs& b, double* s_w, ull* s_nf, double wl, ull nfl) {
  const int lane = (int)__lane_id();
  if (wl != 0.0) atomicAdd(&s_w[lane], wl);
  if (nfl) atomicAdd(&s_nf[lane], nfl);
  __syncthreads();
  if ((int)threadIdx.x < b.nsrc) {
    if (s_w[threadIdx.x] != 0.0) unsafeAtomicAdd(b.wsum + threadIdx.x, s_w[threadIdx.x]);
    if (s_nf[threadIdx.x]) atomicAdd(b.nonfin + threadIdx.x, s_nf[threadIdx.x]);
  }
}

// kBcbV vertices of level r (b.list[0..count)) per wave.  CHILDREN =
// false at the deepest level (no vertex has children there).  Vertices of
// out-degree above kBcbHeavy take a heavy slot (counter *hverts; its row of
// sums zeroed here) and queue their chunks (counter *items).
template <bool CHILDREN>
__device__ __forceinline__ void bcb_light(const BcbArgs& b, const uint64_t* off, const uint32_t* nbr, uint32_t r,
                                          int64_t count, int64_t* items, int64_t* hverts, int64_t wave,
                                          int64_t nwaves, double& wl, ull& nfl) {
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  const ull* lc = lr + b.n;   // level r + 1 (read only when CHILDREN)
  for (int64_t i0 = wave * kBcbV; i0 < count; i0 += nwaves * kBcbV) {
    uint32_t v[kBcbV];
    ull mask[kBcbV];
    uint64_t e0[kBcbV], e1[kBcbV];
    double acc[kBcbV];
#pragma unroll
    for (int t = 0; t < kBcbV; ++t) {
      const bool valid = i0 + t < count;
      v[t] = valid ? b.list[i0 + t] : 0u;
      mask[t] = valid ? lr[v[t]] : 0ull;
      e0[t] = valid && CHILDREN ? off[v[t]] : 0;
      e1[t] = valid && CHILDREN ? off[v[t] + 1] : 0;
      acc[t] = 0.0;
      if (CHILDREN && e1[t] - e0[t] > kBcbHeavy) {
        const uint64_t nch = (e1[t] - e0[t] + kBcbChunk - 1) / kBcbChunk;
        ull h = 0, base = 0;
        if (lane == 0) {
          h = atomicAdd((ull*)hverts, 1ull);
          base = atomicAdd((ull*)items, (ull)nch);
        }
        h = bcb_rl64(h, 0);
        base = bcb_rl64(base, 0);
        if (lane == 0) b.hv[h] = v[t];
        b.acc_h[h * 64 + lane] = 0.0;
        for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = (h << 32) | c;
        mask[t] = 0;        // finished by bcb_heavy_fin
        e1[t] = e0[t];
      }
    }
    if (CHILDREN) gather_rows<kBcbV, false>(nbr, e0, e1, mask, nullptr, lc, b.sig, lane, acc);
#pragma unroll
    for (int t = 0; t < kBcbV; ++t)
      if ((mask[t] >> lane) & 1) bwd_finish(b, v[t], lane, acc[t], wl, nfl);
  }
}

// The chunk items [0, count) that bcb_light queued: per heavy slot sums.
__device__ __forceinline__ void bcb_heavy(const BcbArgs& b, const uint64_t* off, const uint32_t* nbr, uint32_t r,
                                          int64_t count, int64_t wave, int64_t nwaves) {
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  for (int64_t i = wave; i < count; i += nwaves) {
    const ull it = b.items[i];
    const ull h = it >> 32;
    const uint64_t c = it & 0xffffffffull;
    const uint32_t v = b.hv[h];
    const ull mask = lr[v];
    const uint64_t lo = off[v] + c * kBcbChunk;
    const uint64_t hi = std::min<uint64_t>(off[v + 1], lo + kBcbChunk);
    const double s = gather_chunk<false>(nbr, lo, hi, mask, nullptr, lr + b.n, b.sig, lane);
    if (((mask >> lane) & 1) && s != 0.0) unsafeAtomicAdd(b.acc_h + h * 64 + lane, s);
  }
}

// The heavy vertices [0, count) of level r: D from their slot's sums.
__device__ __forceinline__ void bcb_heavy_fin(const BcbArgs& b, uint32_t r, int64_t count, int64_t wave,
                                              int64_t nwaves, double& wl, ull& nfl) {
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  for (int64_t h = wave; h < count; h += nwaves) {
    const uint32_t v = b.hv[h];
    if ((lr[v] >> lane) & 1) bwd_finish(b, v, lane, b.acc_h[h * 64 + lane], wl, nfl);
  }
}

// ---------------------------------------------------------------------------
// A whole batch in the persistent batch kernel: its level loop calls the
// forward phases at every level and runs the backward levels after the
// last one.  Backward counters rotate over three slots (as the level
// outputs of the persistent kernel do).
struct BcbPSlot {
  int64_t count, items, hverts, pad;
};

struct BcPersistHook {
  static constexpr bool kOn = true;
  BcbArgs b;
  const uint64_t* in_off;   // forward gathers: in-edges
  const uint32_t* in_nbr;
  const uint64_t* off;      // backward gathers: out-edges
  const uint32_t* nbr;
  BcbPSlot* slots;          // [3], zero before the launch
};

// Runs the forward of one BC batch (k_bcb_init already launched) and its
// backward as one persistent launch (msbfs.hip); fw gets the forward's
// per-source BFS summaries.
int bfs_batch_bc_persist(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, const BcPersistHook& hk,
                         rg_bfs_summary* fw);

}  // namespace rg
