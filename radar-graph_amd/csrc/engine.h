// Push edgeMap engine for CDNA4 (gfx950): the device side of Ligra's
// edgeMapDenseForward / edgeMapSparse (push-new-orig Ligra/ligra/ligra.h:
// 90-182) with the per-vertex neighbour loop of vertex.h:46-59 and the
// functor's updateAtomic inlined as an "operator" struct.
//
// Work decomposition (one persistent grid per launch):
//   * light kernel: a wavefront takes 64 consecutive frontier entries; each
//     lane owns one vertex; the wave scans the 64 degrees and then walks the
//     concatenated edge ranges 64 edges at a time, each lane finding its
//     source by a 6-step binary search over the scanned degrees in LDS.
//     Consecutive lanes read consecutive column entries (coalesced).
//   * vertices with degree > heavy_deg are not walked there; their edge
//     range is cut into chunks of kHeavyChunk edges that the heavy kernel
//     processes one wavefront per chunk.
//   * hub duplication (HUBDUP / RADAR): destinations that are hubs are
//     updated in a per-workgroup LDS copy instead of with a global atomic;
//     the copy is merged into global state when the workgroup ends.
//
// Operators provide:
//   using SrcT, HubT;
//   static constexpr bool kOutput;   // produces a next frontier
//   SrcT load_src(uint32_t s)        // per-source value, loaded once
//   bool cond(uint32_t d)            // Ligra cond
//   bool update(uint32_t s, SrcT, uint32_t d)   // updateAtomic; true = d joins
//   HubT hub_identity()
//   void hub_update(HubT* lds_slot, uint32_t s, SrcT, uint32_t d)
//   bool hub_flush(uint32_t d, HubT v)          // merge; true = d joins
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace rg {

constexpr int kHeavyChunk = 2048;

enum SrcMode { SRC_ALL = 0, SRC_LIST = 1, SRC_DENSE = 2 };
enum HubMode { HUB_NONE = 0, HUB_BITMAP = 1, HUB_PREFIX = 2 };

struct EdgeArgs {
  const uint64_t* off;
  const uint32_t* nbr;
  int64_t count;          // entries: n for SRC_ALL / SRC_DENSE, list size for SRC_LIST
  const uint32_t* ids;    // SRC_LIST
  const uint8_t* dense;   // SRC_DENSE
  // next frontier (operators with kOutput)
  uint32_t* out_ids;
  int64_t* out_info;      // [0] = count, [1] = sum of out-degrees of the appended vertices
  // heavy work queue
  uint64_t* heavy;        // (vertex << 32 | chunk)
  int64_t* heavy_count;
  int heavy_deg;
  // hubs
  int64_t hub_k;                  // number of hub slots
  const uint32_t* hub_bits;       // HUB_BITMAP: bit per vertex
  const uint32_t* hub_rank;       // HUB_BITMAP: hubs before word w
  const uint32_t* hub_ids;        // HUB_BITMAP: slot -> vertex
};

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (uint64_t(1) << __lane_id()) - 1;
}

// Appends d to the next frontier for every lane with `act`; one atomic per
// wavefront.  Must be called by all lanes of the wave (uniform control flow).
__device__ __forceinline__ void wave_append(const EdgeArgs& a, bool act, uint32_t d) {
  uint64_t mask = __ballot(act);
  if (mask == 0) return;
  uint32_t dd = 0;
  if (act) dd = (uint32_t)(a.off[d + 1] - a.off[d]);
  // Sum of appended degrees (drives the next round's sparse/dense choice).
  uint64_t degsum = dd;
  for (int o = 32; o > 0; o >>= 1) degsum += __shfl_xor(degsum, o);
  int leader = __ffsll((long long)mask) - 1;
  int64_t base = 0;
  if ((int)__lane_id() == leader) {
    base = atomicAdd((unsigned long long*)&a.out_info[0], (unsigned long long)__popcll(mask));
    atomicAdd((unsigned long long*)&a.out_info[1], (unsigned long long)degsum);
  }
  base = __shfl(base, leader);
  if (act) a.out_ids[base + __popcll(mask & lanemask_lt())] = d;
}

template <int HUB>
__device__ __forceinline__ int64_t hub_slot(const EdgeArgs& a, uint32_t d) {
  if (HUB == HUB_PREFIX) return (int64_t)d < a.hub_k ? (int64_t)d : -1;
  if (HUB == HUB_BITMAP) {
    uint32_t w = a.hub_bits[d >> 5];
    uint32_t bit = 1u << (d & 31);
    if (!(w & bit)) return -1;
    return (int64_t)a.hub_rank[d >> 5] + __popc(w & (bit - 1));
  }
  return -1;
}

template <int HUB>
__device__ __forceinline__ uint32_t hub_vertex(const EdgeArgs& a, int64_t slot) {
  return HUB == HUB_BITMAP ? a.hub_ids[slot] : (uint32_t)slot;
}

// One edge: the cond / hub / atomic path.  Returns whether d joins the next
// frontier through a global update.
template <class Op, int HUB>
__device__ __forceinline__ bool do_edge(Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub,
                                        uint32_t s, typename Op::SrcT sv, uint32_t d) {
  if (!op.cond(d)) return false;
  if (HUB != HUB_NONE) {
    int64_t slot = hub_slot<HUB>(a, d);
    if (slot >= 0) {
      op.hub_update(&lds_hub[slot], s, sv, d);
      return false;
    }
  }
  return op.update(s, sv, d);
}

template <class Op, int HUB, int BLOCK>
__device__ __forceinline__ void hub_init(Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub) {
  if (HUB == HUB_NONE) return;
  for (int64_t i = threadIdx.x; i < a.hub_k; i += BLOCK) lds_hub[i] = op.hub_identity();
  __syncthreads();
}

template <class Op, int HUB, int BLOCK>
__device__ __forceinline__ void hub_merge(Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub) {
  if (HUB == HUB_NONE) return;
  __syncthreads();
  const typename Op::HubT ident = op.hub_identity();
  for (int64_t base = 0; base < a.hub_k; base += BLOCK) {
    int64_t slot = base + threadIdx.x;
    bool act = false;
    uint32_t d = 0;
    if (slot < a.hub_k) {
      typename Op::HubT v = lds_hub[slot];
      if (!(v == ident)) {
        d = hub_vertex<HUB>(a, slot);
        act = op.hub_flush(d, v);
      }
    }
    if (Op::kOutput) wave_append(a, act, d);
  }
}

// Light kernel.  Grid-strides over chunks of 64 entries, one chunk per wave.
template <class Op, int MODE, int HUB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void edge_light_kernel(Op op, EdgeArgs a) {
  constexpr int WPB = BLOCK / kWave;
  __shared__ uint32_t s_pre[WPB][kWave + 1];
  __shared__ uint64_t s_start[WPB][kWave];
  __shared__ uint32_t s_src[WPB][kWave];
  __shared__ typename Op::SrcT s_val[WPB][kWave];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  typename Op::HubT* lds_hub = reinterpret_cast<typename Op::HubT*>(dyn_lds);
  if (op.skip()) return;
  hub_init<Op, HUB, BLOCK>(op, a, lds_hub);

  const int lane = (int)__lane_id();
  const int w = threadIdx.x / kWave;
  const int64_t nchunks = (a.count + kWave - 1) / kWave;
  const int64_t wave_id = (int64_t)blockIdx.x * WPB + w;
  const int64_t nwaves = (int64_t)gridDim.x * WPB;

  for (int64_t c = wave_id; c < nchunks; c += nwaves) {
    int64_t i = c * kWave + lane;
    uint32_t v = 0;
    uint32_t deg = 0;
    uint64_t start = 0;
    bool active = i < a.count;
    if (active) {
      if (MODE == SRC_LIST) v = a.ids[i];
      else v = (uint32_t)i;
      if (MODE == SRC_DENSE) active = a.dense[i] != 0;
    }
    if (active) {
      start = a.off[v];
      uint64_t dd = a.off[v + 1] - start;
      if (dd > (uint64_t)a.heavy_deg) {
        int64_t nch = (int64_t)((dd + kHeavyChunk - 1) / kHeavyChunk);
        int64_t base = atomicAdd((unsigned long long*)a.heavy_count, (unsigned long long)nch);
        for (int64_t k = 0; k < nch; ++k) a.heavy[base + k] = ((uint64_t)v << 32) | (uint64_t)k;
        deg = 0;
      } else {
        deg = (uint32_t)dd;
      }
    }
    // Wave inclusive scan of deg.
    uint32_t inc = deg;
    for (int o = 1; o < kWave; o <<= 1) {
      uint32_t t = __shfl_up(inc, o);
      if (lane >= o) inc += t;
    }
    uint32_t total = __shfl(inc, kWave - 1);
    if (total == 0) continue;
    s_pre[w][lane] = inc - deg;
    if (lane == 0) s_pre[w][kWave] = total;
    s_start[w][lane] = start;
    s_src[w][lane] = v;
    if (deg) s_val[w][lane] = op.load_src(v);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
      uint32_t t = t0 + lane;
      bool valid = t < total;
      bool act = false;
      uint32_t d = 0;
      if (valid) {
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (s_pre[w][j + step] <= t) j += step;
        uint64_t e = s_start[w][j] + (t - s_pre[w][j]);
        d = a.nbr[e];
        act = do_edge<Op, HUB>(op, a, lds_hub, s_src[w][j], s_val[w][j], d);
      }
      if (Op::kOutput) wave_append(a, act, d);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  hub_merge<Op, HUB, BLOCK>(op, a, lds_hub);
}

// Heavy kernel: one wavefront per (vertex, chunk) item of the heavy queue.
template <class Op, int HUB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void edge_heavy_kernel(Op op, EdgeArgs a) {
  constexpr int WPB = BLOCK / kWave;
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  typename Op::HubT* lds_hub = reinterpret_cast<typename Op::HubT*>(dyn_lds);
  if (op.skip()) return;
  hub_init<Op, HUB, BLOCK>(op, a, lds_hub);
  const int lane = (int)__lane_id();
  const int64_t items = *a.heavy_count;
  const int64_t wave_id = (int64_t)blockIdx.x * WPB + threadIdx.x / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * WPB;
  for (int64_t it = wave_id; it < items; it += nwaves) {
    uint64_t item = a.heavy[it];
    uint32_t v = (uint32_t)(item >> 32);
    uint64_t k = item & 0xffffffffu;
    uint64_t beg = a.off[v] + k * kHeavyChunk;
    uint64_t end = a.off[v + 1];
    if (end > beg + kHeavyChunk) end = beg + kHeavyChunk;
    typename Op::SrcT sv = op.load_src(v);
    for (uint64_t e0 = beg; e0 < end; e0 += kWave) {
      uint64_t e = e0 + lane;
      bool act = false;
      uint32_t d = 0;
      if (e < end) {
        d = a.nbr[e];
        act = do_edge<Op, HUB>(op, a, lds_hub, v, sv, d);
      }
      if (Op::kOutput) wave_append(a, act, d);
    }
  }
  hub_merge<Op, HUB, BLOCK>(op, a, lds_hub);
}

}  // namespace rg
