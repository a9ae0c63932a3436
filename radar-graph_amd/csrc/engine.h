// Push edgeMap engine for CDNA4 (gfx950): the device side of Ligra's
// edgeMapDenseForward / edgeMapSparse (push-new-orig Ligra/ligra/ligra.h:
// 90-182) with the per-vertex neighbour loop of vertex.h:46-59 and the
// functor's updateAtomic inlined as an "operator" struct.
//
// Work decomposition (one persistent grid per launch):
//   * light phase: a wavefront takes 64 consecutive frontier entries; each
//     lane owns one vertex; the wave scans the 64 degrees and then walks the
//     concatenated edge ranges 64 edges at a time, each lane finding its
//     source by a 6-step binary search over the scanned degrees in LDS.
//     Consecutive lanes read consecutive column entries (coalesced).
//   * heavy phase: vertices with degree > heavy_deg are cut into chunks of
//     kHeavyChunk edges, one wavefront per chunk (lanes stride the chunk).
//     For the all-vertices frontier the chunk list is static and cached per
//     graph, so both phases run in one launch; for list / byte-map frontiers
//     the light phase enqueues chunks and a second launch drains them.
//   * hub duplication (HUBDUP / RADAR): destinations that are hubs are
//     updated in a per-workgroup LDS copy instead of with a global atomic;
//     the copy is merged into global state when the workgroup ends.
//   * next frontier: each wavefront stages activated vertices in its own LDS
//     queue and flushes them in batches (one global atomic per batch), so
//     frontier output does not serialise on a single counter.
//
// Operators provide:
//   using SrcT, HubT;
//   static constexpr bool kOutput;   // produces a next frontier
//   bool skip()                      // whole launch is a no-op (converged)
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
constexpr int kQueue = 256;   // per-wave LDS staging queue (vertices)
// HUBDUP's hub membership filter: a bit per hash of every hub, staged in
// each workgroup's LDS after its hub copy; an arc whose destination's bit is
// clear is no hub and skips the global bitmap lookup.
constexpr int kHubFilterBits = 16;
constexpr int kHubFilterWords = (1 << kHubFilterBits) / 32;

__host__ __device__ inline uint32_t hub_hash(uint32_t d) { return (d * 0x9E3779B1u) >> (32 - kHubFilterBits); }

enum SrcMode { SRC_ALL = 0, SRC_LIST = 1, SRC_DENSE = 2 };
enum HubMode { HUB_NONE = 0, HUB_BITMAP = 1, HUB_PREFIX = 2 };

struct EdgeArgs {
  const uint64_t* off;
  const uint32_t* nbr;
  int64_t count;          // entries: n for SRC_ALL / SRC_DENSE, list size for SRC_LIST
  const int64_t* count_dev; // SRC_LIST: list size read on the device (count is then a bound)
  const uint32_t* ids;    // SRC_LIST
  const uint8_t* dense;   // SRC_DENSE
  // next frontier (operators with kOutput)
  uint32_t* out_ids;
  int64_t* out_info;      // [0] = count, [1] = sum of out-degrees of the appended vertices
  // heavy work: dynamic queue (list / byte-map frontiers) or a static list
  // (all-vertices frontier), items are (vertex << 32 | chunk)
  uint64_t* heavy;
  int64_t* heavy_count;
  const uint64_t* static_heavy;
  int64_t static_heavy_count;
  int heavy_deg;
  const int64_t* heavy_deg_dev;   // SRC_LIST: cut-over read on the device (overrides heavy_deg)
  // hubs.  Slots are ordered by descending in-degree; slots < hub_k live
  // in the workgroup's LDS copy, slots in [hub_k, hub_k2) in the per-XCD
  // copy l2_copies[xcd * hub_k2 + slot] (operators with kL2 only).
  int64_t hub_k;                  // LDS slots
  int64_t hub_k2;                 // LDS + per-XCD slots
  const uint32_t* hub_bits;       // HUB_BITMAP: bit per vertex
  const uint32_t* hub_rank;       // HUB_BITMAP: hubs before word w
  const uint32_t* hub_slot_of;    // HUB_BITMAP: id-order rank -> slot
  const uint32_t* hub_ids;        // HUB_BITMAP: slot -> vertex
  const uint32_t* hub_filter;     // HUB_BITMAP: kHubFilterWords words of hub_hash bits
  void* l2_copies;
  int l2_ncopies;
};

// XCD of the executing wave (HW_REG_XCC_ID); picks the per-XCD hub copy.
__device__ __forceinline__ int xcd_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (uint64_t(1) << __lane_id()) - 1;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wavefront staging queue for the next frontier.  fill is wave-uniform.
struct WaveQueue {
  uint32_t* buf;     // cap entries in LDS
  int fill;
  uint64_t deg;      // this lane's share of the appended out-degrees
  int cap = kQueue;  // a flush (one atomic on the list's shared count) per cap - 64 or more entries
};

__device__ __forceinline__ void wq_flush(WaveQueue& q, const EdgeArgs& a) {
  if (q.fill == 0) return;
  wave_sync_lds();
  unsigned long long base = 0;
  if (__lane_id() == 0) base = atomicAdd((unsigned long long*)&a.out_info[0], (unsigned long long)q.fill);
  base = __shfl(base, 0);
  for (int i = (int)__lane_id(); i < q.fill; i += kWave) a.out_ids[base + i] = q.buf[i];
  wave_sync_lds();
  q.fill = 0;
}

// Appends d for every lane with `act`.  Called by all lanes of the wave.
__device__ __forceinline__ void wq_push(WaveQueue& q, const EdgeArgs& a, bool act, uint32_t d) {
  uint64_t mask = __ballot(act);
  if (mask == 0) return;
  if (act) {
    q.buf[q.fill + __popcll(mask & lanemask_lt())] = d;
    q.deg += a.off[d + 1] - a.off[d];
  }
  q.fill += __popcll(mask);
  if (q.fill > q.cap - kWave) wq_flush(q, a);
}

// For kernels that finalise each vertex once (one vertex per thread): the
// workgroup's joining vertices take one range of the next-frontier list,
// with one atomic for the count and one for the out-degree sum per
// workgroup rather than per wave (thousands of short workgroups flushing
// per wave queue on the same two counters, which serialise).  Every thread
// of the workgroup calls it.
template <int THREADS>
__device__ __forceinline__ void wg_push(const EdgeArgs& a, bool act, uint32_t d) {
  constexpr int NW = THREADS / kWave;
  __shared__ uint32_t s_cnt[NW];
  __shared__ uint64_t s_deg[NW];
  __shared__ unsigned long long s_base;
  const uint64_t mask = __ballot(act);
  const int w = (int)(threadIdx.x / kWave);
  uint64_t deg = act ? a.off[d + 1] - a.off[d] : 0;
  for (int o = kWave / 2; o > 0; o >>= 1) deg += __shfl_xor(deg, o);
  if (__lane_id() == 0) {
    s_cnt[w] = (uint32_t)__popcll(mask);
    s_deg[w] = deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    uint64_t dg = 0;
    for (int i = 0; i < NW; ++i) {
      const uint32_t y = s_cnt[i];
      s_cnt[i] = t;
      t += y;
      dg += s_deg[i];
    }
    s_base = t ? atomicAdd((unsigned long long*)&a.out_info[0], (unsigned long long)t) : 0ull;
    if (dg) atomicAdd((unsigned long long*)&a.out_info[1], (unsigned long long)dg);
  }
  __syncthreads();
  if (act) a.out_ids[s_base + s_cnt[w] + __popcll(mask & lanemask_lt())] = d;
}

// A workgroup's list of joining vertices in LDS (ids[0 .. *n), up to cap),
// appended to the next frontier with one atomic for the count per append
// and one for the out-degree sum at the end, instead of a flush per wave
// queue: kernels with many short workgroups, or levels that join tens of
// millions of vertices, otherwise queue on the list's one count
// (same-address atomics serialise).  *n must be zeroed (and synchronised)
// before the first push; wgl_flush leaves it zeroed.
struct WgList {
  uint32_t* ids;
  uint32_t* n;
  uint32_t cap;
};

// Adds d for the lanes with act; deg accumulates the lane's out-degrees.
// Called by whole waves.
__device__ __forceinline__ void wgl_push(const WgList& L, const EdgeArgs& a, bool act, uint32_t d, uint64_t& deg) {
  const uint64_t mask = __ballot(act);
  if (mask == 0) return;
  uint32_t pos = 0;
  if (__lane_id() == 0) pos = atomicAdd(L.n, (uint32_t)__popcll(mask));
  pos = __shfl(pos, 0);
  if (act) {
    L.ids[pos + __popcll(mask & lanemask_lt())] = d;
    deg += a.off[d + 1] - a.off[d];
  }
}

// Appends the list and empties it when fewer than `room` entries are free
// (a kernel that loops calls it at the end of each step, with room = what
// one step can add).  Every thread of the workgroup calls it.
template <int THREADS>
__device__ __forceinline__ void wgl_drain(const WgList& L, const EdgeArgs& a, uint32_t room) {
  __shared__ unsigned long long s_base;
  __syncthreads();
  const uint32_t n = *L.n;
  if (n + room <= L.cap) return;
  if (threadIdx.x == 0) s_base = atomicAdd((unsigned long long*)&a.out_info[0], (unsigned long long)n);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += THREADS) a.out_ids[s_base + i] = L.ids[i];
  __syncthreads();
  if (threadIdx.x == 0) *L.n = 0;
  __syncthreads();
}

// Appends the rest of the list and the workgroup's out-degree sum (deg is
// zeroed); every thread of the workgroup calls it.
template <int THREADS>
__device__ __forceinline__ void wgl_flush(const WgList& L, const EdgeArgs& a, uint64_t& deg) {
  constexpr int NW = THREADS / kWave;
  __shared__ uint64_t s_deg[NW];
  __shared__ unsigned long long s_base;
  uint64_t dl = deg;
  deg = 0;
  for (int o = kWave / 2; o > 0; o >>= 1) dl += __shfl_xor(dl, o);
  if (__lane_id() == 0) s_deg[threadIdx.x / kWave] = dl;
  __syncthreads();
  const uint32_t n = *L.n;
  if (threadIdx.x == 0) {
    uint64_t dg = 0;
    for (int i = 0; i < NW; ++i) dg += s_deg[i];
    s_base = n ? atomicAdd((unsigned long long*)&a.out_info[0], (unsigned long long)n) : 0ull;
    if (dg) atomicAdd((unsigned long long*)&a.out_info[1], (unsigned long long)dg);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += THREADS) a.out_ids[s_base + i] = L.ids[i];
  __syncthreads();
  if (threadIdx.x == 0) *L.n = 0;
  __syncthreads();
}

__device__ __forceinline__ void wq_finish(WaveQueue& q, const EdgeArgs& a) {
  wq_flush(q, a);
  uint64_t d = q.deg;
  for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
  if (__lane_id() == 0 && d) atomicAdd((unsigned long long*)&a.out_info[1], (unsigned long long)d);
}

template <int HUB>
__device__ __forceinline__ int64_t hub_slot(const EdgeArgs& a, const uint32_t* lds_filter, uint32_t d) {
  if (HUB == HUB_PREFIX) return (int64_t)d < a.hub_k2 ? (int64_t)d : -1;
  if (HUB == HUB_BITMAP) {
    const uint32_t h = hub_hash(d);
    if (!((lds_filter[h >> 5] >> (h & 31)) & 1u)) return -1;
    uint32_t w = a.hub_bits[d >> 5];
    uint32_t bit = 1u << (d & 31);
    if (!(w & bit)) return -1;
    return (int64_t)a.hub_slot_of[a.hub_rank[d >> 5] + __popc(w & (bit - 1))];
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
__device__ __forceinline__ bool do_edge(const Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub,
                                        typename Op::HubT* l2_hub, uint32_t s, typename Op::SrcT sv,
                                        uint32_t d) {
  if (!op.cond(d)) return false;
  if (HUB != HUB_NONE) {
    // A destination the arc cannot change (a claimed BFS vertex, a word
    // that already holds the bits) is dropped before the hub lookup, so the
    // lookup's loads are paid on the arcs that update only.
    if (!op.pre_hub(sv, d)) return false;
    const uint32_t* filt = reinterpret_cast<const uint32_t*>(lds_hub + a.hub_k);
    int64_t slot = hub_slot<HUB>(a, filt, d);
    if (slot >= 0 && slot < a.hub_k) {
      op.hub_update(&lds_hub[slot], s, sv, d);
      return false;
    }
    if (Op::kL2 && slot >= 0) {
      op.l2_update(&l2_hub[slot], sv);
      return false;
    }
  }
  return op.update(s, sv, d);
}

template <class Op, int HUB>
__device__ __forceinline__ typename Op::HubT* l2_copy(const EdgeArgs& a) {
  if (!Op::kL2 || HUB == HUB_NONE || a.hub_k2 <= a.hub_k) return nullptr;
  return reinterpret_cast<typename Op::HubT*>(a.l2_copies) + (int64_t)(xcd_id() % a.l2_ncopies) * a.hub_k2;
}

template <class Op, int HUB, int BLOCK>
__device__ __forceinline__ void hub_init(const Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub) {
  if (HUB == HUB_NONE) return;
  for (int64_t i = threadIdx.x; i < a.hub_k; i += BLOCK) lds_hub[i] = op.hub_identity();
  if (HUB == HUB_BITMAP) {
    uint32_t* filt = reinterpret_cast<uint32_t*>(lds_hub + a.hub_k);
    for (int i = threadIdx.x; i < kHubFilterWords; i += BLOCK) filt[i] = a.hub_filter[i];
  }
  __syncthreads();
}

// Merges the workgroup's LDS copy: into the XCD's L2 copy when there is one
// (the per-XCD merge kernel then covers every slot), else into global state.
template <class Op, int HUB, int BLOCK>
__device__ __forceinline__ void hub_merge(const Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub,
                                          typename Op::HubT* l2_hub, WaveQueue& q) {
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
        if (Op::kL2 && l2_hub) {
          op.l2_update(&l2_hub[slot], v);
        } else {
          d = hub_vertex<HUB>(a, slot);
          act = op.hub_flush(d, v);
        }
      }
    }
    if (Op::kOutput) wq_push(q, a, act, d);
  }
}

// Processes heavy items [first, end) with stride `step`, one per wave.
template <class Op, int HUB>
__device__ __forceinline__ void heavy_items(const Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub,
                                            typename Op::HubT* l2_hub, WaveQueue& q,
                                            const uint64_t* items, int64_t first, int64_t end,
                                            int64_t step) {
  const int lane = (int)__lane_id();
  for (int64_t it = first; it < end; it += step) {
    uint64_t item = items[it];
    uint32_t v = (uint32_t)(item >> 32);
    uint64_t k = item & 0xffffffffu;
    uint64_t beg = a.off[v] + k * kHeavyChunk;
    uint64_t stop = a.off[v + 1];
    if (stop > beg + kHeavyChunk) stop = beg + kHeavyChunk;
    typename Op::SrcT sv = op.load_src(v);
    for (uint64_t e0 = beg; e0 < stop; e0 += kWave) {
      uint64_t e = e0 + lane;
      bool act = false;
      uint32_t d = 0;
      if (e < stop) {
        d = a.nbr[e];
        act = do_edge<Op, HUB>(op, a, lds_hub, l2_hub, v, sv, d);
      }
      if (Op::kOutput) wq_push(q, a, act, d);
    }
  }
}

// LDS of the light phase: per wave, the scanned degrees, edge starts,
// sources and source values of its 64 entries.
template <class Op, int WPB>
struct LightLds {
  uint32_t pre[WPB][kWave + 1];
  uint64_t start[WPB][kWave];
  uint32_t src[WPB][kWave];
  typename Op::SrcT val[WPB][kWave];
};

// Light phase over `count` frontier entries: chunks of 64 entries, one
// chunk per wave (grid-stride over the waves given).  Vertices of degree
// above heavy_deg are queued as chunk items (unless the heavy list is
// static) and skipped here.  Called by whole waves.
// Wave queue entries of the edge kernels: 768 without LDS hub copies, whose
// launches hold little else in LDS (a level that claims millions of vertices
// otherwise flushes its queues tens of thousands of times, each an atomic on
// the list's one count; same-address atomics serialise), kQueue with them
// (the host sizes the hub copies around kQueue).
constexpr int kQueueNoHub = 768;
template <class Op, int HUB>
constexpr int edge_queue() {
  return !Op::kOutput ? 1 : HUB == HUB_NONE ? kQueueNoHub : kQueue;
}

template <class Op, int MODE, int HUB, int WPB>
__device__ __forceinline__ void light_phase(const Op& op, const EdgeArgs& a, typename Op::HubT* lds_hub,
                                            typename Op::HubT* l2_hub, WaveQueue& q, LightLds<Op, WPB>& L,
                                            int64_t count, int64_t wave_id, int64_t nwaves, bool static_heavy,
                                            uint64_t heavy_deg) {
  const int lane = (int)__lane_id();
  const int w = threadIdx.x / kWave;
  const int64_t nchunks = (count + kWave - 1) / kWave;
  for (int64_t c = wave_id; c < nchunks; c += nwaves) {
    int64_t i = c * kWave + lane;
    uint32_t v = 0;
    uint32_t deg = 0;
    uint64_t start = 0;
    bool active = i < count;
    if (active) {
      if (MODE == SRC_LIST) v = a.ids[i];
      else v = (uint32_t)i;
      if (MODE == SRC_DENSE) active = a.dense[i] != 0;
    }
    if (active) {
      start = a.off[v];
      uint64_t dd = a.off[v + 1] - start;
      if (dd > heavy_deg) {
        if (!static_heavy) {
          int64_t nch = (int64_t)((dd + kHeavyChunk - 1) / kHeavyChunk);
          int64_t base = atomicAdd((unsigned long long*)a.heavy_count, (unsigned long long)nch);
          for (int64_t k = 0; k < nch; ++k) a.heavy[base + k] = ((uint64_t)v << 32) | (uint64_t)k;
        }
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
    L.pre[w][lane] = inc - deg;
    L.start[w][lane] = start;
    L.src[w][lane] = v;
    if (deg) L.val[w][lane] = op.load_src(v);
    wave_sync_lds();
    for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
      uint32_t t = t0 + lane;
      bool act = false;
      uint32_t d = 0;
      if (t < total) {
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (L.pre[w][j + step] <= t) j += step;
        uint64_t e = L.start[w][j] + (t - L.pre[w][j]);
        d = a.nbr[e];
        act = do_edge<Op, HUB>(op, a, lds_hub, l2_hub, L.src[w][j], L.val[w][j], d);
      }
      if (Op::kOutput) wq_push(q, a, act, d);
    }
    wave_sync_lds();
  }
}

// Light (+ static heavy) kernel.  Grid-strides over chunks of 64 entries,
// one chunk per wave.
template <class Op, int MODE, int HUB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void edge_light_kernel(Op op, EdgeArgs a) {
  constexpr int WPB = BLOCK / kWave;
  __shared__ LightLds<Op, WPB> L;
  constexpr int QN = edge_queue<Op, HUB>();
  __shared__ uint32_t s_q[Op::kOutput ? WPB : 1][QN];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  typename Op::HubT* lds_hub = reinterpret_cast<typename Op::HubT*>(dyn_lds);
  if (op.skip()) return;
  hub_init<Op, HUB, BLOCK>(op, a, lds_hub);
  typename Op::HubT* l2_hub = l2_copy<Op, HUB>(a);

  const int w = threadIdx.x / kWave;
  WaveQueue q{Op::kOutput ? s_q[w] : nullptr, 0, 0, QN};
  const int64_t count = a.count_dev ? *a.count_dev : a.count;
  const int64_t wave_id = (int64_t)blockIdx.x * WPB + w;
  const int64_t nwaves = (int64_t)gridDim.x * WPB;
  const bool static_heavy = a.static_heavy != nullptr;
  const uint64_t heavy_deg = a.heavy_deg_dev ? (uint64_t)*a.heavy_deg_dev : (uint64_t)a.heavy_deg;
  light_phase<Op, MODE, HUB, WPB>(op, a, lds_hub, l2_hub, q, L, count, wave_id, nwaves, static_heavy, heavy_deg);
  if (static_heavy)
    heavy_items<Op, HUB>(op, a, lds_hub, l2_hub, q, a.static_heavy, wave_id, a.static_heavy_count,
                         nwaves);
  hub_merge<Op, HUB, BLOCK>(op, a, lds_hub, l2_hub, q);
  if (Op::kOutput) wq_finish(q, a);
}

// Heavy kernel: drains the dynamic heavy queue, one wavefront per item.
template <class Op, int HUB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void edge_heavy_kernel(Op op, EdgeArgs a) {
  constexpr int WPB = BLOCK / kWave;
  constexpr int QN = edge_queue<Op, HUB>();
  __shared__ uint32_t s_q[Op::kOutput ? WPB : 1][QN];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  typename Op::HubT* lds_hub = reinterpret_cast<typename Op::HubT*>(dyn_lds);
  if (op.skip()) return;
  const int64_t items = *a.heavy_count;
  if (items == 0) return;
  hub_init<Op, HUB, BLOCK>(op, a, lds_hub);
  typename Op::HubT* l2_hub = l2_copy<Op, HUB>(a);
  const int w = threadIdx.x / kWave;
  WaveQueue q{Op::kOutput ? s_q[w] : nullptr, 0, 0, QN};
  const int64_t wave_id = (int64_t)blockIdx.x * WPB + w;
  const int64_t nwaves = (int64_t)gridDim.x * WPB;
  heavy_items<Op, HUB>(op, a, lds_hub, l2_hub, q, a.heavy, wave_id, items, nwaves);
  hub_merge<Op, HUB, BLOCK>(op, a, lds_hub, l2_hub, q);
  if (Op::kOutput) wq_finish(q, a);
}

// Merges the per-XCD copies of slots [0, hub_k2) into global state and
// clears them (operators with kL2; runs after the edge launch).
template <class Op, int HUB>
__global__ void k_l2_merge(Op op, EdgeArgs a) {
  if (op.skip()) return;
  typename Op::HubT* base = reinterpret_cast<typename Op::HubT*>(a.l2_copies);
  const typename Op::HubT ident = op.hub_identity();
  for (int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; slot < a.hub_k2;
       slot += (int64_t)gridDim.x * blockDim.x) {
    typename Op::HubT s = ident;
    bool any = false;
    for (int x = 0; x < a.l2_ncopies; ++x) {
      typename Op::HubT v = base[(int64_t)x * a.hub_k2 + slot];
      if (!(v == ident)) {
        s = any ? op.hub_combine(s, v) : v;
        any = true;
        base[(int64_t)x * a.hub_k2 + slot] = ident;
      }
    }
    if (any) op.hub_flush(hub_vertex<HUB>(a, slot), s);
  }
}

}  // namespace rg
