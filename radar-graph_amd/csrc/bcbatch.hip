// Batched betweenness centrality: up to 64 independent single-source BC
// queries (BC-manyiter-v2.C:280-398 run with one source each) answered
// together.
//
// Forward: the levels come from the bit-parallel BFS of msbfs.hip (one
// 64-bit word per vertex, source s owning bit s), which hands every level to
// a hook here.  For each vertex d that some source reaches first at level
// L + 1, one wavefront computes the path counts of all those sources at
// once, lane s for source s:
//     sigma_s(d) = sum of sigma_s(u) over the in-neighbours u of d at level L from s
// (BC_F, BC-manyiter-v2.C:189-213, in pull form).  The in-edges are read 64
// at a time, a lane per edge; the edges whose tail is on the current
// frontier of one of d's new sources are then taken one by one with a lane
// per source, each lane reading its own word of the tail's 64-word row of
// path counts: one coalesced 512-byte row per such edge.
//
// Backward, level r from the deepest down to 0 (BC_Back_F and
// BC_Back_Vertex_F, BC-manyiter-v2.C:215-258, in pull form): for each vertex
// v at level r from source s,
//     D_s(v) = 1/sigma_s(v) + sum of D_s(w) over the out-neighbours w at level r + 1 from s
// and the dependency is (D - 1/sigma) / (1/sigma) (BC-manyiter-v2.C:387-389),
// NaN for the vertices the source does not reach (1/0 there).  D replaces
// sigma in place: a (vertex, source) entry is read at one level and written
// at another, so no pass reads an entry that it writes.
//
// Per vertex: sigma / D [64] f64, level [64] u8 (0xFF = unreached) and a
// 64-bit mask of the levels at which some source reaches it.  A vertex of
// degree above kBcbHeavy has its edges split into chunks of kBcbChunk, one
// queue item each, whose sums meet by FP64 atomics.
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

constexpr int kBcbMaxSources = 64;
constexpr int64_t kBcbMaxLevel = 63;   // bits of the level mask
constexpr uint64_t kBcbHeavy = 4096;   // degree above which edges are split into chunks
constexpr uint64_t kBcbChunk = 4096;
constexpr int kBcbThreads = 256;
constexpr int kBcbWaves = kBcbThreads / kWave;

typedef unsigned long long ull;

struct BcbCtl {
  int64_t items;     // heavy chunk items queued at this level
  int64_t hverts;    // heavy vertices of this backward level
  int64_t count;     // vertices of this backward level
  int64_t overflow;  // a source reached a level beyond kBcbMaxLevel
};

struct BcbArgs {
  BcbCtl* ctl;
  int64_t n;
  int nsrc;
  double* sig;           // [n][64] path counts, then D
  uint8_t* lev;          // [n][64] level from each source, 0xFF = unreached
  ull* lvmask;           // [n] levels at which some source reaches the vertex
  ull* items;            // heavy chunk queue: (vertex or heavy slot) << 32 | chunk
  uint32_t* hv;          // backward: heavy vertices of the level
  double* acc_h;         // backward: [heavy slot][64] sums
  uint32_t* list;        // backward: vertices of the level
  ull* bm_cur;           // backward: bitmap of the level's vertices (written by select)
  const ull* bm_child;   // backward: bitmap of level r + 1
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

// Sum of sig[u][lane] over the in-edges [e0, e1) whose tail u is on the
// current frontier of source `lane`, for the lanes in nn (the sources that
// reach the head first at the next level).  Called by whole waves.
__device__ __forceinline__ double fwd_sum(const uint32_t* in_nbr, uint64_t e0, uint64_t e1, ull nn,
                                          const ull* active, const ull* front, const double* sig, int lane) {
  double acc = 0.0;
  for (uint64_t e = e0; e < e1; e += kWave) {
    const uint64_t k = e + lane;
    uint32_t u = 0;
    ull f = 0;
    if (k < e1) {
      u = in_nbr[k];
      if ((active[u >> 6] >> (u & 63)) & 1) f = front[u] & nn;
    }
    ull hit = __ballot(f != 0);
    while (hit) {
      // Four tails per step, their row loads issued together.
      uint32_t uu[4];
      bool on[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bool have = hit != 0;
        const int j = have ? __ffsll((long long)hit) - 1 : 0;
        hit &= hit - 1;
        const ull fj = bcb_rl64(f, j);
        uu[t] = bcb_rl32(u, j);
        on[t] = have && ((fj >> lane) & 1);
      }
      double x[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) x[t] = on[t] ? sig[(size_t)uu[t] * 64 + lane] : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) acc += x[t];
    }
  }
  return acc;
}

// Sum of D[w][lane] over the out-edges [e0, e1) whose head w is at level rc
// from source `lane`, for the lanes where `mine`.  bm: vertices at level rc
// from some source.  Called by whole waves.
__device__ __forceinline__ double bwd_sum(const uint32_t* nbr, uint64_t e0, uint64_t e1, uint32_t rc, const ull* bm,
                                          const uint8_t* lev, const double* D, bool mine, int lane) {
  double acc = 0.0;
  for (uint64_t e = e0; e < e1; e += kWave) {
    const uint64_t k = e + lane;
    uint32_t w = 0;
    bool c = false;
    if (k < e1) {
      w = nbr[k];
      c = (bm[w >> 6] >> (w & 63)) & 1;
    }
    ull hit = __ballot(c);
    while (hit) {
      uint32_t ww[4];
      bool have[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        have[t] = hit != 0;
        const int j = have[t] ? __ffsll((long long)hit) - 1 : 0;
        hit &= hit - 1;
        ww[t] = bcb_rl32(w, j);
      }
      uint32_t l[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) l[t] = (have[t] && mine) ? lev[(size_t)ww[t] * 64 + lane] : 0xFFu;
      double x[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) x[t] = l[t] == rc ? D[(size_t)ww[t] * 64 + lane] : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) acc += x[t];
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Forward (level hook).

__global__ void k_bcb_reset(BcbCtl* c) {
  c->items = 0;
  c->hverts = 0;
  c->count = 0;
}

// Wave per vertex of the next frontier: its level bytes and level-mask bit,
// then its path counts (vertices of in-degree above kBcbHeavy: zeroed here,
// summed by k_bcf_heavy's chunk items).
__global__ __launch_bounds__(kBcbThreads) void k_bcf_light(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone) return;
  const int64_t count = *v.ncount;
  const int64_t L1 = *v.level + 1;
  if (L1 > kBcbMaxLevel) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && count > 0) b.ctl->overflow = 1;
    return;
  }
  const int lane = (int)__lane_id();
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
    const uint32_t d = v.list_nxt[i];
    const ull nn = v.front_nxt[d];
    const bool mine = (nn >> lane) & 1;
    const size_t row = (size_t)d * 64 + lane;
    if (mine) b.lev[row] = (uint8_t)L1;
    if (lane == 0) b.lvmask[d] |= 1ull << L1;
    const uint64_t e0 = in_off[d], e1 = in_off[d + 1];
    if (e1 - e0 > kBcbHeavy) {
      if (mine) b.sig[row] = 0.0;
      const uint64_t nch = (e1 - e0 + kBcbChunk - 1) / kBcbChunk;
      ull base = 0;
      if (lane == 0) base = atomicAdd((ull*)&b.ctl->items, (ull)nch);
      base = bcb_rl64(base, 0);
      for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = ((ull)d << 32) | c;
      continue;
    }
    const double s = fwd_sum(in_nbr, e0, e1, nn, v.active_cur, v.front_cur, b.sig, lane);
    if (mine) b.sig[row] = s;
  }
}

__global__ __launch_bounds__(kBcbThreads) void k_bcf_heavy(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone || *v.level + 1 > kBcbMaxLevel) return;
  const int64_t count = b.ctl->items;
  const int lane = (int)__lane_id();
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
    const ull it = b.items[i];
    const uint32_t d = (uint32_t)(it >> 32);
    const uint64_t c = it & 0xffffffffull;
    const ull nn = v.front_nxt[d];
    const uint64_t lo = in_off[d] + c * kBcbChunk;
    const uint64_t hi = std::min<uint64_t>(in_off[d + 1], lo + kBcbChunk);
    const double s = fwd_sum(in_nbr, lo, hi, nn, v.active_cur, v.front_cur, b.sig, lane);
    if (((nn >> lane) & 1) && s != 0.0) unsafeAtomicAdd(b.sig + (size_t)d * 64 + lane, s);
  }
}

// Sources at level 0; per-source sums cleared.
__global__ void k_bcb_init(BcbArgs b, const uint32_t* sources) {
  for (int s = threadIdx.x; s < b.nsrc; s += blockDim.x) {
    const uint32_t v = sources[s];
    b.sig[(size_t)v * 64 + s] = 1.0;
    b.lev[(size_t)v * 64 + s] = 0;
    atomicOr(b.lvmask + v, 1ull);
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

// Vertices at level r from some source: the list and the bitmap.
__global__ __launch_bounds__(kBcbThreads) void k_bcb_select(BcbArgs b, uint32_t r) {
  const int lane = (int)__lane_id();
  for (int64_t base = bcb_wave() * kWave; base < b.n; base += bcb_waves() * kWave) {
    const int64_t v = base + lane;
    const bool at = v < b.n && ((b.lvmask[v] >> r) & 1);
    const ull word = __ballot(at);
    if (lane == 0) b.bm_cur[base >> 6] = word;
    if (word) {
      ull pos = 0;
      if (lane == 0) pos = atomicAdd((ull*)&b.ctl->count, (ull)__popcll(word));
      pos = bcb_rl64(pos, 0);
      if (at) b.list[pos + __popcll(word & ((1ull << lane) - 1))] = (uint32_t)v;
    }
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

// Wave per vertex of level r.  CHILDREN = false at the deepest level (no
// vertex has children there).  Vertices of out-degree above kBcbHeavy take a
// heavy slot (its row of sums zeroed here) and queue their chunks.
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
  double wl = 0.0;
  ull nfl = 0;
  const int64_t count = b.ctl->count;
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
    const uint32_t v = b.list[i];
    const bool mine = b.lev[(size_t)v * 64 + lane] == r;
    double acc = 0.0;
    if (CHILDREN) {
      const uint64_t e0 = off[v], e1 = off[v + 1];
      if (e1 - e0 > kBcbHeavy) {
        const uint64_t nch = (e1 - e0 + kBcbChunk - 1) / kBcbChunk;
        ull h = 0, base = 0;
        if (lane == 0) {
          h = atomicAdd((ull*)&b.ctl->hverts, 1ull);
          base = atomicAdd((ull*)&b.ctl->items, (ull)nch);
        }
        h = bcb_rl64(h, 0);
        base = bcb_rl64(base, 0);
        if (lane == 0) b.hv[h] = v;
        b.acc_h[h * 64 + lane] = 0.0;
        for (uint64_t c = lane; c < nch; c += kWave) b.items[base + c] = (h << 32) | c;
        continue;
      }
      acc = bwd_sum(nbr, e0, e1, r + 1, b.bm_child, b.lev, b.sig, mine, lane);
    }
    if (mine) bwd_finish(b, v, lane, acc, wl, nfl);
  }
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

__global__ __launch_bounds__(kBcbThreads) void k_bcb_heavy(BcbArgs b, const uint64_t* off, const uint32_t* nbr,
                                                           uint32_t r) {
  const int64_t count = b.ctl->items;
  const int lane = (int)__lane_id();
  for (int64_t i = bcb_wave(); i < count; i += bcb_waves()) {
    const ull it = b.items[i];
    const ull h = it >> 32;
    const uint64_t c = it & 0xffffffffull;
    const uint32_t v = b.hv[h];
    const bool mine = b.lev[(size_t)v * 64 + lane] == r;
    const uint64_t lo = off[v] + c * kBcbChunk;
    const uint64_t hi = std::min<uint64_t>(off[v + 1], lo + kBcbChunk);
    const double s = bwd_sum(nbr, lo, hi, r + 1, b.bm_child, b.lev, b.sig, mine, lane);
    if (mine && s != 0.0) unsafeAtomicAdd(b.acc_h + h * 64 + lane, s);
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
  double wl = 0.0;
  ull nfl = 0;
  const int64_t count = b.ctl->hverts;
  for (int64_t h = bcb_wave(); h < count; h += bcb_waves()) {
    const uint32_t v = b.hv[h];
    if (b.lev[(size_t)v * 64 + lane] == r) bwd_finish(b, v, lane, b.acc_h[h * 64 + lane], wl, nfl);
  }
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

// ---------------------------------------------------------------------------
// Host side.

struct BcbState {
  DevBuf sig, lev, lvmask, items, hv, acc_h, list, bm[2], ctl, stats, srcs;
  int64_t n = -1, m = -1;
};

void bcb_state_free(void* p) { delete static_cast<BcbState*>(p); }

namespace {

int64_t heavy_cap(int64_t m) { return m / (int64_t)kBcbHeavy + 1; }
int64_t item_cap(int64_t m) { return 2 * heavy_cap(m) + 64; }

int bcb_prepare(rg_ctx* ctx, BcbState& S, int64_t n, int64_t m) {
  if (S.n == n && S.m == m) return RG_OK;
  for (DevBuf* b : {&S.sig, &S.lev, &S.lvmask, &S.items, &S.hv, &S.acc_h, &S.list, &S.bm[0], &S.bm[1]}) b->release();
  const size_t words = (size_t)(n / 64 + 1);
  RG_TRY(S.sig.ensure(sizeof(double) * 64 * (size_t)n));
  RG_TRY(S.lev.ensure((size_t)64 * n));
  RG_TRY(S.lvmask.ensure(sizeof(ull) * (size_t)n));
  RG_TRY(S.items.ensure(sizeof(ull) * (size_t)item_cap(m)));
  RG_TRY(S.hv.ensure(sizeof(uint32_t) * (size_t)heavy_cap(m)));
  RG_TRY(S.acc_h.ensure(sizeof(double) * 64 * (size_t)heavy_cap(m)));
  RG_TRY(S.list.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(n, 1)));
  RG_TRY(S.bm[0].ensure(sizeof(ull) * words));
  RG_TRY(S.bm[1].ensure(sizeof(ull) * words));
  RG_TRY(S.ctl.ensure(sizeof(BcbCtl)));
  RG_TRY(S.stats.ensure(sizeof(double) * 2 * kBcbMaxSources));
  RG_TRY(S.srcs.ensure(sizeof(uint32_t) * kBcbMaxSources));
  S.n = n;
  S.m = m;
  (void)ctx;
  return RG_OK;
}

struct ForwardHook : MsLevelHook {
  rg_ctx* ctx;
  BcbArgs b;
  const uint64_t* in_off;
  const uint32_t* in_nbr;
  int grid;
  int level(const MsLevelView& v) override {
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_bcb_reset, 1, 1, 0, st, b.ctl);
    hipLaunchKernelGGL(k_bcf_light, grid, kBcbThreads, 0, st, b, v, in_off, in_nbr);
    hipLaunchKernelGGL(k_bcf_heavy, grid, kBcbThreads, 0, st, b, v, in_off, in_nbr);
    RG_HIP(hipGetLastError());
    return RG_OK;
  }
};

// Sources whose levels exceed the level mask: one single-source run each
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
  return n * (64 * 8 + 64 + 8 + 4) + 2 * (n / 64 + 1) * 8 + item_cap(m) * 8 + heavy_cap(m) * (4 + 64 * 8) + 4096;
}

int bc_batch(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_out,
             rg_bc_summary* out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (nsrc < 1 || nsrc > kBcbMaxSources) return set_error(RG_EINVAL, "rg_bc_batch: between 1 and 64 sources");
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
  RG_TRY(bcb_prepare(ctx, *S, n, g->m));
  DevBuf dep;
  double* deps_dev = nullptr;
  if (deps_out) {
    RG_TRY(dep.ensure(sizeof(double) * (size_t)nsrc * n));
    deps_dev = dep.as<double>();
  }
  hipStream_t st = ctx->stream;
  const int grid = ctx->num_cus * 8;
  EventTimer tm(st);
  if (deps_dev) RG_TRY(fill<double>(ctx, deps_dev, (int64_t)nsrc * n, std::nan("")));
  RG_TRY(fill<ull>(ctx, S->lev.as<ull>(), n * 8, ~0ull));
  RG_TRY(fill<ull>(ctx, S->lvmask.as<ull>(), n, 0ull));
  RG_HIP(hipMemcpyAsync(S->srcs.p, sources, sizeof(uint32_t) * nsrc, hipMemcpyHostToDevice, st));
  BcbArgs b{};
  b.ctl = S->ctl.as<BcbCtl>();
  b.n = n;
  b.nsrc = nsrc;
  b.sig = S->sig.as<double>();
  b.lev = S->lev.as<uint8_t>();
  b.lvmask = S->lvmask.as<ull>();
  b.items = S->items.as<ull>();
  b.hv = S->hv.as<uint32_t>();
  b.acc_h = S->acc_h.as<double>();
  b.list = S->list.as<uint32_t>();
  b.wsum = S->stats.as<double>();
  b.nonfin = reinterpret_cast<ull*>(S->stats.as<double>() + kBcbMaxSources);
  b.deps = deps_dev;
  hipLaunchKernelGGL(k_bcb_init, 1, 64, 0, st, b, S->srcs.as<uint32_t>());
  RG_HIP(hipGetLastError());

  // Forward: the batch BFS with the path-count hook at every level.
  ForwardHook hook;
  hook.ctx = ctx;
  hook.b = b;
  hook.in_off = g->in_off;
  hook.in_nbr = g->in_nbr;
  hook.grid = grid;
  std::vector<rg_bfs_summary> fw(nsrc);
  RG_TRY(bfs_batch_hooked(g, sources, nsrc, o, &hook, fw.data()));
  int64_t overflow = 0;
  RG_TRY(ctx_read_count(ctx, &b.ctl->overflow, &overflow, 1));
  if (overflow) {
    RG_TRY(bc_batch_single(g, sources, nsrc, o, deps_out, fw, out));
    if (ms_out) *ms_out = tm.stop_ms();
    return RG_OK;
  }
  int64_t maxd = 0;
  for (const auto& x : fw) maxd = std::max(maxd, x.iters - 1);

  // Backward, deepest level first.
  const int gsel = (int)std::max<int64_t>(1, std::min<int64_t>((n / 64 + kBcbWaves) / kBcbWaves, grid));
  for (int64_t r = maxd; r >= 0; --r) {
    b.bm_cur = S->bm[r & 1].as<ull>();
    b.bm_child = S->bm[(r + 1) & 1].as<ull>();
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
  std::vector<double> h(2 * kBcbMaxSources);
  RG_TRY(ctx_copy(ctx, h.data(), S->stats.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  const double ms = tm.stop_ms();
  for (int s = 0; s < nsrc; ++s) {
    uint64_t nf = 0;
    std::memcpy(&nf, &h[kBcbMaxSources + s], sizeof(nf));
    out[s].rounds = fw[s].iters;
    out[s].reached = fw[s].reached;
    out[s].wsum = h[s];
    out[s].nonfinite = (int64_t)nf + (n - fw[s].reached);
  }
  if (ms_out) *ms_out = ms;
  if (deps_out)
    RG_TRY(ctx_copy(ctx, deps_out, deps_dev, sizeof(double) * (size_t)nsrc * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

}  // namespace rg

extern "C" int rg_bc_batch(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
                           rg_bc_summary* out, double* ms_out) {
  if (!g) return rg::set_error(RG_EINVAL, "rg_bc_batch: null graph");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::bc_batch(g, sources, nsources, o, deps_out, out, ms_out);
}
