// Host side of the push engine: the per-call configuration (Engine), the
// launch helpers that size the persistent grids, small fill kernels and an
// event timer.  Shared by the app loops and the multi-source BFS.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "device.h"
#include "engine.h"

namespace rg {

template <class T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <class T>
__global__ void k_set(T* p, int64_t i, T v) {
  p[i] = v;
}

inline int grid_for(rg_ctx* ctx, int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  int64_t cap = (int64_t)ctx->num_cus * 16;
  return (int)std::max<int64_t>(1, std::min(g, cap));
}

template <class T>
inline int fill(rg_ctx* ctx, T* p, int64_t n, T v) {
  if (n <= 0) return RG_OK;
  hipLaunchKernelGGL(k_fill<T>, grid_for(ctx, n), 256, 0, ctx->stream, p, n, v);
  RG_HIP(hipGetLastError());
  return RG_OK;
}

// ---------------------------------------------------------------------------
// Engine configuration and the edgeMap dispatcher.

// Capacity of the heavy queue in items: at most m / heavy_deg heavy
// vertices, each split in chunks of kHeavyChunk arcs.
inline int64_t heavy_queue_items(int64_t m, int heavy_deg) {
  return m / kHeavyChunk + m / std::max(1, heavy_deg) + 64;
}

struct Engine {
  rg_ctx* ctx;
  rg_graph* g;
  int64_t threshold;   // Ligra: m/20 by default (ligra.h:269)
  int block;           // 256 or 1024
  int heavy_deg;
  int hub_mode;        // HUB_NONE / HUB_BITMAP / HUB_PREFIX
  int64_t hub_k;       // LDS slots
  int64_t hub_k2;      // LDS + per-XCD slots
  void* l2_copies;
  int l2_ncopies;
  uint64_t* heavy;     // heavy work queue (chunk items)
  int64_t heavy_cap = 0;   // its capacity in items (0: never lower the cut-over)
  int64_t* heavy_count;
  // List / byte-map frontiers with fewer arcs (|F| + out-degree sum) run
  // without hub copies: their few workgroups would spend as long setting up
  // and merging the copies as on the arcs.  0: always with copies.
  int64_t hub_min_arcs = 0;
  const HubMeta* hm = nullptr;   // HUB_BITMAP: the hub metadata in use
};

template <class Op, int MODE, int HUB, int BLOCK>
static int launch_edge(const Engine& E, const Op& op, const EdgeArgs& a, size_t hub_bytes) {
  constexpr int WPB = BLOCK / kWave;
  auto kl = edge_light_kernel<Op, MODE, HUB, BLOCK>;
  auto kh = edge_heavy_kernel<Op, HUB, BLOCK>;
  // Largest dynamic LDS size set on these kernels so far (contexts may
  // launch from several host threads).
  static std::atomic<size_t> attr_set{0};
  if (hub_bytes > 0 && hub_bytes > attr_set.load()) {
    RG_HIP(hipFuncSetAttribute((const void*)kl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hub_bytes));
    RG_HIP(hipFuncSetAttribute((const void*)kh, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hub_bytes));
    size_t cur = attr_set.load();
    while (hub_bytes > cur && !attr_set.compare_exchange_weak(cur, hub_bytes)) {
    }
  }
  const int64_t nchunks = (a.count + kWave - 1) / kWave;
  int64_t want = (nchunks + WPB - 1) / WPB;
  // Persistent grid: enough workgroups to fill every CU; with hub copies, as
  // many as the LDS allows per CU (each extra copy costs one merge).
  int64_t per_cu;
  if (HUB != HUB_NONE) {
    size_t static_lds = WPB * ((kWave + 1) * 4 + kWave * (8 + 4 + 8));
    per_cu = std::max<int64_t>(1, (int64_t)(E.ctx->lds_per_cu / (static_lds + hub_bytes)));
    per_cu = std::min<int64_t>(per_cu, std::max(1, 32 / WPB));
  } else {
    per_cu = std::max(1, 64 / WPB);
  }
  int64_t cap = (int64_t)E.ctx->num_cus * per_cu;
  int grid_l = (int)std::max<int64_t>(1, std::min(want, cap));
  if (a.static_heavy) grid_l = (int)cap;   // the heavy phase needs the full grid
  hipEvent_t* ev = timing_begin(E.ctx);
  hipLaunchKernelGGL(kl, grid_l, BLOCK, hub_bytes, E.ctx->stream, op, a);
  RG_HIP(hipGetLastError());
  if (ev) RG_HIP(hipEventRecord(ev[1], E.ctx->stream));
  if (!a.static_heavy) {
    // The heavy queue is usually short or empty: a smaller persistent grid
    // (its waves stride over the items) keeps an empty launch cheap.
    const int grid_h = (int)std::max<int64_t>(1, std::min<int64_t>(cap, (int64_t)E.ctx->num_cus * 4));
    hipLaunchKernelGGL(kh, grid_h, BLOCK, hub_bytes, E.ctx->stream, op, a);
    RG_HIP(hipGetLastError());
  }
  if (Op::kL2 && HUB != HUB_NONE && a.hub_k2 > a.hub_k) {
    int64_t span = a.hub_k2;
    int g = (int)std::max<int64_t>(1, std::min<int64_t>((span + 255) / 256, (int64_t)E.ctx->num_cus * 8));
    hipLaunchKernelGGL((k_l2_merge<Op, HUB>), g, 256, 0, E.ctx->stream, op, a);
    RG_HIP(hipGetLastError());
  }
  return RG_OK;
}

template <class Op, int MODE, int HUB>
static int launch_block(const Engine& E, const Op& op, const EdgeArgs& a, size_t hub_bytes) {
  if (E.block >= 1024) return launch_edge<Op, MODE, HUB, 1024>(E, op, a, hub_bytes);
  return launch_edge<Op, MODE, HUB, 256>(E, op, a, hub_bytes);
}

template <class Op, int MODE>
static int launch_hub(const Engine& E, const Op& op, EdgeArgs a) {
  size_t hub_bytes = (size_t)a.hub_k * sizeof(typename Op::HubT);
  switch (E.hub_mode) {
    case HUB_BITMAP:   // the copy, then the membership filter
      return launch_block<Op, MODE, HUB_BITMAP>(E, op, a, hub_bytes + sizeof(uint32_t) * kHubFilterWords);
    case HUB_PREFIX: return launch_block<Op, MODE, HUB_PREFIX>(E, op, a, hub_bytes);
    default: a.hub_k = a.hub_k2 = 0; return launch_block<Op, MODE, HUB_NONE>(E, op, a, 0);
  }
}

// Resolves variant/opts into an Engine.  elem = bytes of one hub copy entry;
// wgs_per_cu = workgroups per CU the default LDS hub copy leaves room for
// (1 for the accumulating operators, bound by memory-side atomics, whose
// copies pay at any size; 4 for the traversal operators, bound by latency,
// which need the occupancy).  l2_tier: the operator takes the per-XCD hub
// copies (Op::kL2; hubs past the LDS tier are otherwise plain vertices, and
// the hub metadata covers the LDS tier alone).  heavy / heavy_cap /
// heavy_count (all given or none): the heavy queue and its counter in caller
// memory instead of the context pools; only the push variant (no hub
// copies) then runs without allocating.
int make_engine(rg_graph* g, const rg_opts* o, size_t elem, Engine& E, int wgs_per_cu = 1, bool l2_tier = false,
                uint64_t* heavy = nullptr, int64_t heavy_cap = 0, int64_t* heavy_count = nullptr);

// ---------------------------------------------------------------------------
// Timing helper (device time of a region on the context stream).
struct EventTimer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit EventTimer(hipStream_t st) : s(st) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  double stop_ms() {
    float ms = 0.f;
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
  }
  ~EventTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace rg
