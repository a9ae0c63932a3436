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

#include "bcbatch.h"
#include "device.h"
#include "engine.h"
#include "engine_host.h"
#include "queries.h"

namespace rg {

__device__ __forceinline__ int64_t bcb_wave() {
  return (int64_t)blockIdx.x * kBcbWaves + threadIdx.x / kWave;
}
__device__ __forceinline__ int64_t bcb_waves() { return (int64_t)gridDim.x * kBcbWaves; }

// ---------------------------------------------------------------------------
// Forward (level hook): one kernel per phase.

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
  bcf_words(b, L1, v.front_nxt, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, (int64_t)gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(kBcbThreads) void k_bcf_light(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone || *v.level + 1 >= b.lcap) return;
  bcf_light(b, *v.ncount, v.list_nxt, v.front_nxt, v.front_cur, v.active_cur, in_off, in_nbr, &b.ctl->items,
            bcb_wave(), bcb_waves());
}

__global__ __launch_bounds__(kBcbThreads) void k_bcf_heavy(BcbArgs b, MsLevelView v, const uint64_t* in_off,
                                                           const uint32_t* in_nbr) {
  if (*v.mode == kMsDone || *v.level + 1 >= b.lcap) return;
  bcf_heavy(b, b.ctl->items, v.front_nxt, v.front_cur, v.active_cur, in_off, in_nbr, bcb_wave(), bcb_waves());
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
// Backward: one kernel per phase.

__global__ __launch_bounds__(kBcbThreads) void k_bcb_select(BcbArgs b, uint32_t r) {
  bcb_select(b, r, &b.ctl->count, bcb_wave(), bcb_waves());
}

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
  double wl = 0.0;
  ull nfl = 0;
  bcb_light<CHILDREN>(b, off, nbr, r, b.ctl->count, &b.ctl->items, &b.ctl->hverts, bcb_wave(), bcb_waves(), wl,
                      nfl);
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

__global__ __launch_bounds__(kBcbThreads) void k_bcb_heavy(BcbArgs b, const uint64_t* off, const uint32_t* nbr,
                                                           uint32_t r) {
  bcb_heavy(b, off, nbr, r, b.ctl->items, bcb_wave(), bcb_waves());
}

__global__ __launch_bounds__(kBcbThreads) void k_bcb_heavy_fin(BcbArgs b, uint32_t r) {
  __shared__ double s_w[64];
  __shared__ ull s_nf[64];
  if (threadIdx.x < 64) {
    s_w[threadIdx.x] = 0.0;
    s_nf[threadIdx.x] = 0;
  }
  __syncthreads();
  double wl = 0.0;
  ull nfl = 0;
  bcb_heavy_fin(b, r, b.ctl->hverts, bcb_wave(), bcb_waves(), wl, nfl);
  bcb_flush(b, s_w, s_nf, wl, nfl);
}

// ---------------------------------------------------------------------------
// Host side.

struct BcbState {
  DevBuf sig, lw, items, hv, acc_h, list, ctl, stats, srcs, pslots;
  int64_t n = -1, m = -1;
  int lcap = 0;
};

void bcb_state_free(void* p) { delete static_cast<BcbState*>(p); }

int64_t bcb_state_bytes(const void* p) {
  const BcbState* S = static_cast<const BcbState*>(p);
  if (!S) return 0;
  int64_t b = 0;
  for (const DevBuf* d :
       {&S->sig, &S->lw, &S->items, &S->hv, &S->acc_h, &S->list, &S->ctl, &S->stats, &S->srcs, &S->pslots})
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
    RG_TRY(S.pslots.ensure(sizeof(BcbPSlot) * 3));
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

  if (o && o->batch_launch == 2) {
    // Forward and backward in one persistent launch.  Only on request: the
    // gathers of BC are bound by load latency, and inside the persistent
    // grid they run at the occupancy its dense traversal phase allows (two
    // 512-thread workgroups per CU) instead of the per-level kernels' eight
    // 256-thread ones -- symmetric R-MAT of 2^22 / 2^25 vertices, batch of
    // 64: 36.8 / 265.9 ms against 30.7 / 255.7 ms per-level.
    RG_TRY(fill<int64_t>(ctx, S.pslots.as<int64_t>(), (int64_t)(sizeof(BcbPSlot) * 3 / sizeof(int64_t)), 0));
    const BcPersistHook hk{b, g->in_off, g->in_nbr, g->off, g->nbr, S.pslots.as<BcbPSlot>()};
    RG_TRY(bfs_batch_bc_persist(g, sources, nsrc, o, hk, fw.data()));
    int64_t ovf = 0;
    RG_TRY(ctx_read_count(ctx, &b.ctl->overflow, &ovf, 1));
    *overflow = ovf != 0;
    return RG_OK;
  }

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
// time; batch_launch 1 and 2 keep every batch bit-parallel (2: each batch
// as one persistent launch).
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
    if (k < kBcbSingleBelow && !(o && o->batch_launch != 0)) {
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
