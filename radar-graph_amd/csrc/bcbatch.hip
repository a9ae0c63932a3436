// Batched betweenness centrality: up to 64 independent single-source BC
// queries (BC-manyiter-v2.C:124-241 run with one source each) answered
// together.
//
// Forward: the levels come from the bit-parallel BFS of msbfs.hip (one
// 64-bit word per vertex, source s owning bit s), which hands every level to
// a hook here.  For each vertex d that some source reaches first at level
// L + 1, one wavefront computes the path counts of all those sources at
// once, lane s for source s:
//     sigma_s(d) = sum of sigma_s(u) over the in-neighbours u of d at level L from s
// (BC_F, BC-manyiter-v2.C:32-56, in pull form).  The in-edges are read 64
// at a time, a lane per edge; the edges whose tail is on the current
// frontier of one of d's new sources are then taken one by one with a lane
// per source, each lane reading its own word of the tail's 64-word row of
// path counts: one coalesced 512-byte row per such edge.
//
// Backward, level r from the deepest down to 0 (BC_Back_F and
// BC_Back_Vertex_F, BC-manyiter-v2.C:58-101, in pull form): for each vertex
// v at level r from source s,
//     D_s(v) = 1/sigma_s(v) + sum of D_s(w) over the out-neighbours w at level r + 1 from s
// and the dependency is (D - 1/sigma) / (1/sigma) (BC-manyiter-v2.C:230-232),
// NaN for the vertices the source does not reach (1/0 there).  The same
// gather as the forward one, over out-edges: a lane per edge tests the
// head's level-(r + 1) word against v's level-r word, then a lane per source
// reads the heads' rows.  D replaces sigma in place: a (vertex, source)
// entry is read at one level and written at another, so no pass reads an
// entry that it writes.
//
// Per vertex: sigma / D [64] f64, and per level L a 64-bit word of the
// sources for which the vertex is at level L (the bit-parallel BFS's
// frontier word of that level, kept).  A vertex of degree above kBcbHeavy
// has its edges split into chunks of kBcbChunk, one queue item each, whose
// sums meet by FP64 atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.h"
#include "engine.h"
#include "engine_host.h"
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
constexpr int kBcbThreads = 256;
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

__device__ __forceinline__ int64_t bcb_wave() {
  return (int64_t)blockIdx.x * kBcbWaves + threadIdx.x / kWave;
}
__device__ __forceinline__ int64_t bcb_waves() { return (int64_t)gridDim.x * kBcbWaves; }

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
// Forward (level hook).

__global__ void k_bcb_reset(BcbCtl* c) {
  c->items = 0;
  c->hverts = 0;
  c->count = 0;
}

// The next level's words: lw[L + 1] = front_nxt (every vertex written).
__global__ __launch_bounds__(kBcbThreads) void k_bcf_words(BcbArgs b, MsLevelView v) {
  if (*v.mode == kMsDone) return;
  const int64_t L1 = *v.level + 1;
  if (L1 >= b.lcap) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *v.ncount > 0) b.ctl->overflow = 1;
    return;
  }
  ull* dst = b.lw + L1 * b.n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < b.n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v.front_nxt[i];
}

// kBcbV vertices of the next frontier per wave: their path counts
// (vertices of in-degree above kBcbHeavy: zeroed here, summed by
// k_bcf_heavy's chunk items).

__global__ __launch_bounds__(kBcbThreads) void k_bcf_light(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone || *v.level + 1 >= b.lcap) return;
  const int64_t count = *v.ncount;
  const int lane = (int)__lane_id();
  for (int64_t i0 = bcb_wave() * kBcbV; i0 < count; i0 += bcb_waves() * kBcbV) {
    uint32_t d[kBcbV];
    ull nn[kBcbV];
    uint64_t e0[kBcbV], e1[kBcbV];
    double acc[kBcbV];
#pragma unroll
    for (int t = 0; t < kBcbV; ++t) {
      const bool valid = i0 + t < count;
      d[t] = valid ? v.list_nxt[i0 + t] : 0u;
      nn[t] = valid ? v.front_nxt[d[t]] : 0ull;
      e0[t] = valid ? in_off[d[t]] : 0;
      e1[t] = valid ? in_off[d[t] + 1] : 0;
      acc[t] = 0.0;
      if (e1[t] - e0[t] > kBcbHeavy) {
        if ((nn[t] >> lane) & 1) b.sig[(size_t)d[t] * 64 + lane] = 0.0;
        const uint64_t nch = (e1[t] - e0[t] + kBcbChunk - 1) / kBcbChunk;
        ull base = 0;
        if (lane == 0) base = atomicAdd((ull*)&b.ctl->items, (ull)nch);
        base = bcb_rl64(base, 0);
        for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = ((ull)d[t] << 32) | c;
        nn[t] = 0;          // nothing more to write for this vertex here
        e1[t] = e0[t];
      }
    }
    gather_rows<kBcbV, true>(in_nbr, e0, e1, nn, v.active_cur, v.front_cur, b.sig, lane, acc);
#pragma unroll
    for (int t = 0; t < kBcbV; ++t)
      if ((nn[t] >> lane) & 1) b.sig[(size_t)d[t] * 64 + lane] = acc[t];
  }
}

__global__ __launch_bounds__(kBcbThreads) void k_bcf_heavy(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone || *v.level + 1 >= b.lcap) return;
  const int64_t count = b.ctl->items;
  const int lane = (int)__lane_id();
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
    const ull it = b.items[i];
    const uint32_t d = (uint32_t)(it >> 32);
    const uint64_t c = it & 0xffffffffull;
    const ull nn = v.front_nxt[d];
    const uint64_t lo = in_off[d] + c * kBcbChunk;
    const uint64_t hi = std::min<uint64_t>(in_off[d + 1], lo + kBcbChunk);
    const double s = gather_chunk<true>(in_nbr, lo, hi, nn, v.active_cur, v.front_cur, b.sig, lane);
    if (((nn >> lane) & 1) && s != 0.0) unsafeAtomicAdd(b.sig + (size_t)d * 64 + lane, s);
  }
}

// Sources at level 0 (lw[0] cleared before); per-source sums cleared.
__global__ void k_bcb_init(BcbArgs b, const uint32_t* sources) {
  for (int s = threadIdx.x; s < b.nsrc; s += blockDim.x) {
    const uint32_t v = sources[s];
    b.sig[(size_t)v * 64 + s] = 1.0;
    atomicOr(b.lw + v, 1ull << s);
  }
  for (int s = threadIdx.x; s < kBcbMaxSources; s += blockDim.x) {
    b.wsum[s] = 0.0;
    b.nonfin[s] = 0;
  }
  if (threadIdx.x == 0) {
    b.ctl->items = 0;
    b.ctl->hverts = 0;
    b.ctl->count = 0;
    b.ctl->overflow = 0;
  }
}

// ---------------------------------------------------------------------------
// Backward.

// Vertices at level r from some source.  Each wave scans a contiguous
// range of 64-vertex words twice: counting, one reservation in the list,
// writing.
__global__ __launch_bounds__(kBcbThreads) void k_bcb_select(BcbArgs b, uint32_t r) {
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  const int64_t nwords = (b.n + kWave - 1) / kWave;
  const int64_t per = (nwords + bcb_waves() - 1) / bcb_waves();
  const int64_t w0 = bcb_wave() * per, w1 = std::min<int64_t>(w0 + per, nwords);
  ull total = 0;
  for (int64_t w = w0; w < w1; ++w) {
    const int64_t v = w * kWave + lane;
    total += __popcll(__ballot(v < b.n && lr[v] != 0));
  }
  if (!total) return;
  ull pos = 0;
  if (lane == 0) pos = atomicAdd((ull*)&b.ctl->count, total);
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

// Per-lane sums of a workgroup into the per-source totals.
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

// kBcbV vertices of level r per wave.  CHILDREN = false at the deepest
// level (no vertex has children there).  Vertices of out-degree above
// kBcbHeavy take a heavy slot (its row of sums zeroed here) and queue their
// chunks.
template <bool CHILDREN>
__global__ __launch_bounds__(kBcbThreads) void k_bcb_light(BcbArgs b, const uint64_t* off, const uint32_t* nbr,
                                                           uint32_t r) {
  __shared__ double s_w[64];
  __shared__ ull s_nf[64];
  if (threadIdx.x < 64) {
    s_w[threadIdx.x] = 0.0;
    s_nf[threadIdx.x] = 0;
  }
  __syncthreads();
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  const ull* lc = lr + b.n;   // level r + 1 (read only when CHILDREN)
  double wl = 0.0;
  ull nfl = 0;
  const int64_t count = b.ctl->count;
  for (int64_t i0 = bcb_wave() * kBcbV; i0 < count; i0 += bcb_waves() * kBcbV) {
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
          h = atomicAdd((ull*)&b.ctl->hverts, 1ull);
          base = atomicAdd((ull*)&b.ctl->items, (ull)nch);
        }
        h = bcb_rl64(h, 0);
        base = bcb_rl64(base, 0);
        if (lane == 0) b.hv[h] = v[t];
        b.acc_h[h * 64 + lane] = 0.0;
        for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = (h << 32) | c;
        mask[t] = 0;        // finished by k_bcb_heavy_fin
        e1[t] = e0[t];
      }
    }
    if (CHILDREN) gather_rows<kBcbV, false>(nbr, e0, e1, mask, nullptr, lc, b.sig, lane, acc);
#pragma unroll
    for (int t = 0; t < kBcbV; ++t)
      if ((mask[t] >> lane) & 1) bwd_finish(b, v[t], lane, acc[t], wl, nfl);
  }
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

__global__ __launch_bounds__(kBcbThreads) void k_bcb_heavy(BcbArgs b, const uint64_t* off, const uint32_t* nbr,
                                                           uint32_t r) {
  const int64_t count = b.ctl->items;
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
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

__global__ __launch_bounds__(kBcbThreads) void k_bcb_heavy_fin(BcbArgs b, uint32_t r) {
  __shared__ double s_w[64];
  __shared__ ull s_nf[64];
  if (threadIdx.x < 64) {
    s_w[threadIdx.x] = 0.0;
    s_nf[threadIdx.x] = 0;
  }
  __syncthreads();
  const int lane = (int)__lane_id();
  const ull* lr = b.lw + (int64_t)r * b.n;
  double wl = 0.0;
  ull nfl = 0;
  const int64_t count = b.ctl->hverts;
  for (int64_t h = bcb_wave(); h < count; h += bcb_waves()) {
    const uint32_t v = b.hv[h];
    if ((lr[v] >> lane) & 1) bwd_finish(b, v, lane, b.acc_h[h * 64 + lane], wl, nfl);
  }
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

// ---------------------------------------------------------------------------
// Host side.

struct BcbState {
  DevBuf sig, lw, items, hv, acc_h, list, ctl, stats, srcs;
  int64_t n = -1, m = -1;
  int lcap = 0;
};

void bcb_state_free(void* p) { delete static_cast<BcbState*>(p); }

int64_t bcb_state_bytes(const void* p) {
  const BcbState* S = static_cast<const BcbState*>(p);
  if (!S) return 0;
  int64_t b = 0;
  for (const DevBuf* d : {&S->sig, &S->lw, &S->items, &S->hv, &S->acc_h, &S->list, &S->ctl, &S->stats, &S->srcs})
    b += (int64_t)d->bytes;
  return b;
}

namespace {

int64_t heavy_cap(int64_t m) { return m / (int64_t)kBcbHeavy + 1; }
int64_t item_cap(int64_t m) { return 2 * heavy_cap(m) + 64; }

int bcb_prepare(BcbState& S, int64_t n, int64_t m, int lcap) {
  if (S.n != n || S.m != m) {
    for (DevBuf* b : {&S.sig, &S.lw, &S.items, &S.hv, &S.acc_h, &S.list}) b->release();
    RG_TRY(S.sig.ensure(sizeof(double) * 64 * (size_t)n));
    RG_TRY(S.items.ensure(sizeof(ull) * (size_t)item_cap(m)));
    RG_TRY(S.hv.ensure(sizeof(uint32_t) * (size_t)heavy_cap(m)));
    RG_TRY(S.acc_h.ensure(sizeof(double) * 64 * (size_t)heavy_cap(m)));
    RG_TRY(S.list.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(n, 1)));
    RG_TRY(S.ctl.ensure(sizeof(BcbCtl)));
    RG_TRY(S.stats.ensure(sizeof(double) * 2 * kBcbMaxSources));
    RG_TRY(S.srcs.ensure(sizeof(uint32_t) * kBcbMaxSources));
    S.n = n;
    S.m = m;
    S.lcap = 0;
  }
  if (lcap > S.lcap) {
    S.lw.release();
    RG_TRY(S.lw.ensure(sizeof(ull) * (size_t)lcap * (size_t)std::max<int64_t>(n, 1)));
    S.lcap = lcap;
  }
  return RG_OK;
}

struct ForwardHook : MsLevelHook {
  rg_ctx* ctx;
  BcbArgs b;
  const uint64_t* in_off;
  const uint32_t* in_nbr;
  int grid, gwords;
  int level(const MsLevelView& v) override {
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_bcb_reset, 1, 1, 0, st, b.ctl);
    hipLaunchKernelGGL(k_bcf_words, gwords, kBcbThreads, 0, st, b, v);
    hipLaunchKernelGGL(k_bcf_light, grid, kBcbThreads, 0, st, b, v, in_off, in_nbr);
    hipLaunchKernelGGL(k_bcf_heavy, grid, kBcbThreads, 0, st, b, v, in_off, in_nbr);
    RG_HIP(hipGetLastError());
    return RG_OK;
  }
};

// One batch: forward on the bit-parallel BFS with the path-count hook, then
// the backward levels.  *overflow: a source went deeper than the level
// words allocated (nothing of the backward ran).
int bcb_run(rg_graph* g, BcbState& S, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_dev,
            std::vector<rg_bfs_summary>& fw, bool* overflow) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  hipStream_t st = ctx->stream;
  const int grid = ctx->num_cus * 8;
  if (deps_dev) RG_TRY(fill<double>(ctx, deps_dev, (int64_t)nsrc * n, std::nan("")));
  RG_TRY(fill<ull>(ctx, S.lw.as<ull>(), n, 0ull));   // level 0's words
  RG_HIP(hipMemcpyAsync(S.srcs.p, sources, sizeof(uint32_t) * nsrc, hipMemcpyHostToDevice, st));
  BcbArgs b{};
  b.ctl = S.ctl.as<BcbCtl>();
  b.n = n;
  b.nsrc = nsrc;
  b.lcap = S.lcap;
  b.sig = S.sig.as<double>();
  b.lw = S.lw.as<ull>();
  b.items = S.items.as<ull>();
  b.hv = S.hv.as<uint32_t>();
  b.acc_h = S.acc_h.as<double>();
  b.list = S.list.as<uint32_t>();
  b.wsum = S.stats.as<double>();
  b.nonfin = reinterpret_cast<ull*>(S.stats.as<double>() + kBcbMaxSources);
  b.deps = deps_dev;
  hipLaunchKernelGGL(k_bcb_init, 1, 64, 0, st, b, S.srcs.as<uint32_t>());
  RG_HIP(hipGetLastError());

  ForwardHook hook;
  hook.ctx = ctx;
  hook.b = b;
  hook.in_off = g->in_off;
  hook.in_nbr = g->in_nbr;
  hook.grid = grid;
  hook.gwords = grid_for(ctx, n, kBcbThreads);
  RG_TRY(bfs_batch_hooked(g, sources, nsrc, o, &hook, fw.data()));
  int64_t ovf = 0;
  RG_TRY(ctx_read_count(ctx, &b.ctl->overflow, &ovf, 1));
  *overflow = ovf != 0;
  if (*overflow) return RG_OK;
  int64_t maxd = 0;
  for (const auto& x : fw) maxd = std::max(maxd, x.iters - 1);

  // Backward, deepest level first.
  const int gsel = (int)std::max<int64_t>(1, std::min<int64_t>((n / 64 + kBcbWaves) / kBcbWaves, grid));
  for (int64_t r = maxd; r >= 0; --r) {
    hipLaunchKernelGGL(k_bcb_reset, 1, 1, 0, st, b.ctl);
    hipLaunchKernelGGL(k_bcb_select, gsel, kBcbThreads, 0, st, b, (uint32_t)r);
    if (r == maxd) {
      hipLaunchKernelGGL(k_bcb_light<false>, grid, kBcbThreads, 0, st, b, g->off, g->nbr, (uint32_t)r);
    } else {
      hipLaunchKernelGGL(k_bcb_light<true>, grid, kBcbThreads, 0, st, b, g->off, g->nbr, (uint32_t)r);
      hipLaunchKernelGGL(k_bcb_heavy, grid, kBcbThreads, 0, st, b, g->off, g->nbr, (uint32_t)r);
      hipLaunchKernelGGL(k_bcb_heavy_fin, grid, kBcbThreads, 0, st, b, (uint32_t)r);
    }
    RG_HIP(hipGetLastError());
  }
  return RG_OK;
}

// Sources deeper than the level words can hold: one single-source run each
// (rg_bc with that source, the orientation restored after it).
int bc_batch_single(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_out,
                    const std::vector<rg_bfs_summary>& fw, rg_bc_summary* out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  for (int s = 0; s < nsrc; ++s) {
    const bool tr = g->transposed;
    int64_t rounds = 0;
    double ms = 0.0;
    RG_TRY(bc_push(g, &sources[s], 1, o, deps_out ? deps_out + (size_t)s * n : nullptr, &rounds, &ms));
    if (g->transposed != tr) RG_TRY(rg_graph_transpose(g));
    double w = 0.0;
    int64_t nf = 0;
    RG_TRY(fp_summary(ctx, ctx->last_result, n, &w, &nf));
    out[s].rounds = rounds;
    out[s].reached = fw[s].reached;
    out[s].wsum = w;
    out[s].nonfinite = nf;
  }
  return RG_OK;
}

}  // namespace

int64_t bc_batch_bytes(int64_t n, int64_t m) {
  return n * (64 * 8 + 8 * kBcbLevelsInit + 4) + item_cap(m) * 8 + heavy_cap(m) * (4 + 64 * 8) + 4096;
}

namespace {

// One batch of 1..64 sources; deps_dev (nullable, nsrc * n) on the device,
// deps_host the caller's rows, which the single-source fallback writes
// itself (*on_host set).
int bc_batch64(rg_graph* g, BcbState& S, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_dev,
               double* deps_host, rg_bc_summary* out, bool* on_host) {
  *on_host = false;
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  RG_TRY(bcb_prepare(S, n, g->m, kBcbLevelsInit));
  std::vector<rg_bfs_summary> fw(nsrc);
  bool overflow = false;
  RG_TRY(bcb_run(g, S, sources, nsrc, o, deps_dev, fw, &overflow));
  if (overflow) {
    // The forward told the depth: grow the level words and run again, or,
    // beyond kBcbMaxLevels levels, answer each source on its own.
    int64_t levels = 0;
    for (const auto& x : fw) levels = std::max(levels, x.iters);
    if (levels <= kBcbMaxLevels) {
      RG_TRY(bcb_prepare(S, n, g->m, (int)levels));
      RG_TRY(bcb_run(g, S, sources, nsrc, o, deps_dev, fw, &overflow));
    }
    if (overflow) {
      *on_host = true;
      return bc_batch_single(g, sources, nsrc, o, deps_host, fw, out);
    }
  }
  std::vector<double> h(2 * kBcbMaxSources);
  RG_TRY(ctx_copy(ctx, h.data(), S.stats.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  for (int s = 0; s < nsrc; ++s) {
    uint64_t nf = 0;
    std::memcpy(&nf, &h[kBcbMaxSources + s], sizeof(nf));
    out[s].rounds = fw[s].iters;
    out[s].reached = fw[s].reached;
    out[s].wsum = h[s];
    out[s].nonfinite = (int64_t)nf + (n - fw[s].reached);
  }
  return RG_OK;
}

}  // namespace

int bc_batch_reserve(rg_graph* g) {
  rg_ctx* ctx = g->ctx;
  BcbState* S = static_cast<BcbState*>(ctx->bcb);
  if (!S) {
    S = new BcbState;
    ctx->bcb = S;
  }
  RG_TRY(bcb_prepare(*S, g->n, g->m, kBcbLevelsInit));
  return bfs_batch_reserve(g, kBcbMaxSources);
}

// Fewer than kBcbSingleBelow sources as single-source runs (rg_bc with one
// source each, the orientation restored after it): the vertices a source
// reaches are those whose dependency is finite.
int bc_small_single(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_out,
                    rg_bc_summary* out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  for (int s = 0; s < nsrc; ++s) {
    const bool tr = g->transposed;
    int64_t rounds = 0;
    double ms = 0.0;
    RG_TRY(bc_push(g, &sources[s], 1, o, deps_out ? deps_out + (size_t)s * n : nullptr, &rounds, &ms));
    if (g->transposed != tr) RG_TRY(rg_graph_transpose(g));
    double w = 0.0;
    int64_t nf = 0;
    RG_TRY(fp_summary(ctx, ctx->last_result, n, &w, &nf));
    out[s].rounds = rounds;
    out[s].reached = n - nf;
    out[s].wsum = w;
    out[s].nonfinite = nf;
  }
  return RG_OK;
}

// 1..256 sources: batches of 64 in order, the last one uneven (a batch
// costs far less than 64 single runs at any size, so full batches first).
// With the library's choice of launch form (rg_opts::batch_launch 0), a
// last batch of fewer than kBcbSingleBelow sources runs one source at a
// time; batch_launch 1 keeps every batch bit-parallel.
int bc_batch(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_out,
             rg_bc_summary* out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (nsrc < 1 || nsrc > kBcbMaxBatch) return set_error(RG_EINVAL, "rg_bc_batch: between 1 and 256 sources");
  if (!sources || !out) return set_error(RG_EINVAL, "rg_bc_batch: null argument");
  for (int i = 0; i < nsrc; ++i)
    if ((int64_t)sources[i] >= n) return set_error(RG_EINVAL, "rg_bc_batch: source out of range");
  if (n >= ((int64_t)1 << 32) - 1) return set_error(RG_EUNSUPPORTED, "rg_bc_batch: n >= 2^32 - 1");
  RG_TRY(bfs_batch_prepare(g));   // the forward levels' blocked layout, outside the timed region
  BcbState* S = static_cast<BcbState*>(ctx->bcb);
  if (!S) {
    S = new BcbState;
    ctx->bcb = S;
  }
  const int first = std::min(nsrc, kBcbMaxSources);
  DevBuf dep;
  double* deps_dev = nullptr;
  if (deps_out) {
    RG_TRY(dep.ensure(sizeof(double) * (size_t)first * n));
    deps_dev = dep.as<double>();
  }
  double ms = 0.0;
  for (int c0 = 0; c0 < nsrc; c0 += kBcbMaxSources) {
    const int k = std::min(kBcbMaxSources, nsrc - c0);
    double* host = deps_out ? deps_out + (size_t)c0 * n : nullptr;
    EventTimer tm(ctx->stream);
    bool on_host = false;
    if (k < kBcbSingleBelow && !(o && o->batch_launch == 1)) {
      on_host = true;
      RG_TRY(bc_small_single(g, sources + c0, k, o, host, out + c0));
    } else {
      RG_TRY(bc_batch64(g, *S, sources + c0, k, o, deps_dev, host, out + c0, &on_host));
    }
    ms += tm.stop_ms();
    if (deps_out && !on_host)
      RG_TRY(ctx_copy(ctx, host, deps_dev, sizeof(double) * (size_t)k * n, hipMemcpyDeviceToHost));
  }
  if (ms_out) *ms_out = ms;
  return RG_OK;
}

}  // namespace rg

extern "C" int rg_bc_batch(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
                           rg_bc_summary* out, double* ms_out) {
  if (!g) return rg::set_error(RG_EINVAL, "rg_bc_batch: null graph");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::bc_batch(g, sources, nsources, o, deps_out, out, ms_out);
}
