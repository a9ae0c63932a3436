// Triangle counting: GAP's OrderedCount (push-new-orig GAP/src/tc_local.cc:
// 48-86).  A triangle u > v > w is counted once, from its largest vertex u:
// for every neighbour v < u of u, the neighbours w < v of v that are also
// neighbours of u.  With sorted lists, N+(x) = {y in N(x) : y < x} is a
// prefix of N(x), so the count for arc (u, v) is |N+(u) & N+(v)|.
// Each triangle adds 1 to the per-vertex counts of u, v and w
// (tc_local.cc:67-69).  On a degree-relabelled graph (ids by descending
// degree, RelabelByDegree) the prefixes are the "oriented low-to-high
// degree" lists and stay short.
//
// Mapping: one wavefront per vertex u.  With per-vertex counts its lanes take
// the arcs (u, v) of N+(u) and each merges N+(u) with N+(v).  For the total
// alone, arcs (u, v) with v below a split are counted from v's side
// (k_tc_bits): a workgroup puts N+(v) in an LDS bitmap and every u of N-(v)
// (neighbours above v) tests its neighbours below v against it -- the
// shorter list scanned, the longer one probed in one step; on a
// degree-relabelled R-MAT of 2^22 vertices this is 4x fewer checks than
// scanning N+(v) for each u.  The remaining arcs are counted from u's side
// with the (v, w) work flattened over the lanes (k_tc_flat).  The split is
// the first vertex of degree below kTcBitsDeg (at most 2^19, the bitmap's
// range), found on the device: each v of k_tc_bits costs a workgroup-wide
// set-up, which pays for high-degree v only (degree-relabelled R-MAT of
// 2^22 vertices: 38.7 ms with the split at 2^19, 29.3 ms at 2^15, where
// degrees cross ~950).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "engine.h"
#include "engine_host.h"

namespace rg {
namespace {

// Length of the prefix of the sorted list [b, e) holding values < x.
__device__ __forceinline__ uint64_t prefix_below(const uint32_t* nbr, uint64_t b, uint64_t e, uint32_t x) {
  uint64_t lo = b, hi = e;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (nbr[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo - b;
}

__global__ __launch_bounds__(256) void k_tc(const uint64_t* off, const uint32_t* nbr, int64_t n,
                                            unsigned long long* total, unsigned long long* tri) {
  const int lane = (int)__lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t u = wid; u < n; u += nw) {
    const uint64_t ub = off[u], ue = off[u + 1];
    uint64_t du = 0;
    if (lane == 0) du = prefix_below(nbr, ub, ue, (uint32_t)u);
    du = __shfl(du, 0);
    unsigned long long mine = 0;
    for (uint64_t i = lane; i < du; i += kWave) {
      const uint32_t v = nbr[ub + i];
      const uint64_t vb = off[v];
      const uint64_t dv = prefix_below(nbr, vb, off[v + 1], v);
      // Merge N+(u)[0, du) with N+(v)[0, dv): every w of N+(v) is < v, so
      // only the part of N+(u) below v can match.
      uint64_t a = 0, b = 0;
      unsigned long long c = 0;
      while (a < du && b < dv) {
        uint32_t x = nbr[ub + a], y = nbr[vb + b];
        if (x < y) {
          a++;
        } else if (y < x) {
          b++;
        } else {
          c++;
          if (tri) atomicAdd(&tri[y], 1ull);
          a++;
          b++;
        }
      }
      if (tri && c) atomicAdd(&tri[v], c);
      mine += c;
    }
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
    if (lane == 0 && mine) {
      atomicAdd(total, mine);
      if (tri) atomicAdd(&tri[u], mine);
    }
  }
}

// Totals only: the same count with the (v, w) work of a vertex u spread
// evenly over the wavefront.  N+(u) (ascending) is staged in the wave's LDS
// when it fits; for each group of 64 v = N+(u)[i] the wave scans |N+(v)|,
// then every lane takes one w of the concatenated N+(v) lists and
// binary-searches it in N+(u)[0, i) (a common neighbour w < v must sit below
// v's position in N+(u)).  N+(u)[0] is skipped: nothing in N+(u) lies below it.
constexpr int kTcLds = 1024;

template <bool LDS>
__device__ __forceinline__ bool tc_find(const uint32_t* nu_lds, const uint32_t* nu_glb, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t y = LDS ? nu_lds[mid] : nu_glb[mid];
    if (y < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < len && (LDS ? nu_lds[lo] : nu_glb[lo]) == x;
}

template <bool LDS>
__device__ __forceinline__ unsigned long long tc_vertex(const uint64_t* off, const uint32_t* nbr, uint64_t ub,
                                                        uint32_t du, const uint32_t* nu, uint32_t* pre,
                                                        uint64_t* vbeg, uint32_t* vidx, uint32_t istart) {
  const int lane = (int)__lane_id();
  const uint32_t* nu_glb = nbr + ub;
  unsigned long long c = 0;
  for (uint32_t i0 = istart; i0 < du; i0 += kWave) {
    const uint32_t i = i0 + lane;
    uint32_t dv = 0;
    uint64_t vb = 0;
    if (i < du) {
      const uint32_t v = LDS ? nu[i] : nu_glb[i];
      vb = off[v];
      dv = (uint32_t)prefix_below(nbr, vb, off[v + 1], v);
    }
    uint32_t inc = dv;
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o);
      if (lane >= o) inc += t;
    }
    const uint32_t total = __shfl(inc, kWave - 1);
    pre[lane] = inc - dv;
    vbeg[lane] = vb;
    vidx[lane] = i;
    if (lane == 0) pre[kWave] = total;
    wave_sync_lds();
    for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
      const uint32_t t = t0 + lane;
      if (t < total) {
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (pre[j + step] <= t) j += step;
        const uint32_t w = nbr[vbeg[j] + (t - pre[j])];
        if (tc_find<LDS>(nu, nu_glb, vidx[j], w)) ++c;
      }
    }
    wave_sync_lds();
  }
  return c;
}

// Arcs (u, v) with v >= vlo only (the others are k_tc_bits').
__global__ __launch_bounds__(256) void k_tc_flat(const uint64_t* off, const uint32_t* nbr, int64_t n,
                                                 unsigned long long* total, const unsigned long long* split) {
  const uint32_t vlo = (uint32_t)*split;
  constexpr int WPB = 256 / kWave;
  __shared__ uint32_t s_nu[WPB][kTcLds];
  __shared__ uint32_t s_pre[WPB][kWave + 1];
  __shared__ uint64_t s_vb[WPB][kWave];
  __shared__ uint32_t s_vi[WPB][kWave];
  const int lane = (int)__lane_id();
  const int w = threadIdx.x / kWave;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  unsigned long long c = 0;
  for (int64_t u = wid; u < n; u += nw) {
    const uint64_t ub = off[u], ue = off[u + 1];
    uint64_t du = 0;
    if (lane == 0) du = prefix_below(nbr, ub, ue, (uint32_t)u);
    du = __shfl(du, 0);
    if (du < 2) continue;
    // N+(u)[0] has nothing below it; the arcs to v < vlo are k_tc_bits'.
    uint64_t is = 0;
    if (lane == 0) is = prefix_below(nbr, ub, ub + du, vlo);
    const uint32_t istart = (uint32_t)std::max<uint64_t>(1, __shfl(is, 0));
    if (istart >= du) continue;
    if (du <= (uint64_t)kTcLds) {
      for (uint64_t i = lane; i < du; i += kWave) s_nu[w][i] = nbr[ub + i];
      wave_sync_lds();
      c += tc_vertex<true>(off, nbr, ub, (uint32_t)du, s_nu[w], s_pre[w], s_vb[w], s_vi[w], istart);
    } else {
      c += tc_vertex<false>(off, nbr, ub, (uint32_t)du, nullptr, s_pre[w], s_vb[w], s_vi[w], istart);
    }
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0 && c) atomicAdd(total, c);
}

constexpr int kTcBitsLog = 19;
constexpr int64_t kTcBitsV = int64_t(1) << kTcBitsLog;   // bitmap range: vertices v counted by k_tc_bits
constexpr int kTcBitsThreads = 1024;
constexpr int kTcBitsDeg = 512;   // default degree from which v is counted by k_tc_bits

// split = the first vertex below vcap of degree < dmin (vcap if none).
__global__ void k_tc_split(const uint64_t* off, int64_t vcap, uint64_t dmin, unsigned long long* split) {
  unsigned long long first = ~0ull;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < vcap; v += (int64_t)gridDim.x * blockDim.x)
    if (off[v + 1] - off[v] < dmin) {
      first = (unsigned long long)v;
      break;
    }
  for (int o = 32; o > 0; o >>= 1) first = min(first, (unsigned long long)__shfl_xor(first, o));
  if (__lane_id() == 0 && first != ~0ull) atomicMin(split, first);
}

// Persistent workgroups taking vertices v < vmax from a counter: N+(v) as
// bits of an LDS bitmap over [0, 2^19) (set, used, then cleared again), and
// for each u of N-(v), 64 per wave, the neighbours w < v of u (the prefix of
// N(u) below v) tested against it, the (u, w) work flattened over the lanes.
__global__ __launch_bounds__(kTcBitsThreads) void k_tc_bits(const uint64_t* off, const uint32_t* nbr,
                                                            const unsigned long long* split,
                                                            unsigned long long* next_v, unsigned long long* total) {
  const int64_t vmax = (int64_t)*split;
  constexpr int WPB = kTcBitsThreads / kWave;
  __shared__ uint32_t bits[kTcBitsV / 32];
  __shared__ uint32_t s_pre[WPB][kWave + 1];
  __shared__ uint64_t s_ub[WPB][kWave];
  __shared__ int64_t s_v;
  __shared__ uint64_t s_b, s_e, s_dp;
  const int lane = (int)__lane_id();
  const int w = threadIdx.x / kWave;
  for (int i = threadIdx.x; i < (int)(kTcBitsV / 32); i += kTcBitsThreads) bits[i] = 0;
  unsigned long long c = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const int64_t v = (int64_t)atomicAdd(next_v, 1ull);
      s_v = v;
      if (v < vmax) {
        const uint64_t b = off[v], e = off[v + 1];
        s_b = b;
        s_e = e;
        s_dp = prefix_below(nbr, b, e, (uint32_t)v);
      }
    }
    __syncthreads();
    const int64_t v = s_v;
    if (v >= vmax) break;
    const uint64_t b = s_b, e = s_e, dp = s_dp;
    if (dp == 0 || b + dp >= e) continue;   // no neighbour below v, or none above
    for (uint64_t i = threadIdx.x; i < dp; i += kTcBitsThreads) {
      const uint32_t x = nbr[b + i];
      atomicOr(&bits[x >> 5], 1u << (x & 31));
    }
    __syncthreads();
    for (uint64_t j0 = b + dp + (uint64_t)w * kWave; j0 < e; j0 += (uint64_t)WPB * kWave) {
      const uint64_t j = j0 + lane;
      uint32_t r = 0;
      uint64_t ub = 0;
      if (j < e) {
        const uint32_t u = nbr[j];
        ub = off[u];
        r = (uint32_t)prefix_below(nbr, ub, off[u + 1], (uint32_t)v);
      }
      uint32_t inc = r;
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
      }
      const uint32_t tot = __shfl(inc, kWave - 1);
      s_pre[w][lane] = inc - r;
      s_ub[w][lane] = ub;
      if (lane == 0) s_pre[w][kWave] = tot;
      wave_sync_lds();
      for (uint32_t t0 = 0; t0 < tot; t0 += kWave) {
        const uint32_t t = t0 + lane;
        if (t < tot) {
          int q = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1)
            if (s_pre[w][q + step] <= t) q += step;
          const uint32_t x = nbr[s_ub[w][q] + (t - s_pre[w][q])];
          c += (bits[x >> 5] >> (x & 31)) & 1u;
        }
      }
      wave_sync_lds();
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < dp; i += kTcBitsThreads) bits[nbr[b + i] >> 5] = 0;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0 && c) atomicAdd(total, c);
}

}  // namespace
}  // namespace rg

using rg::set_error;

extern "C" int rg_tc(rg_graph* g, const rg_opts* o, uint64_t* total_out, uint64_t* per_vertex,
                     double* ms_out) {
  if (!g || !total_out) return set_error(RG_EINVAL, "rg_tc: null argument");
  if (!g->symmetric)
    return set_error(RG_EUNSUPPORTED, "Input graph is directed but tc requires undirected");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = g->n;
  rg::DevBuf tri;
  unsigned long long* d_total = reinterpret_cast<unsigned long long*>(ctx->counts.as<int64_t>() + 20);
  RG_HIP(hipMemsetAsync(d_total, 0, sizeof(uint64_t), s));
  if (per_vertex) {
    RG_TRY(tri.ensure(sizeof(uint64_t) * n));
    RG_HIP(hipMemsetAsync(tri.p, 0, sizeof(uint64_t) * n, s));
  }
  hipEvent_t a, b;
  RG_HIP(hipEventCreate(&a));
  RG_HIP(hipEventCreate(&b));
  RG_HIP(hipEventRecord(a, s));
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, (int64_t)ctx->num_cus * 32));
  if (per_vertex)
    hipLaunchKernelGGL(rg::k_tc, grid, 256, 0, s, g->off, g->nbr, n, d_total, tri.as<unsigned long long>());
  else {
    const int64_t vcap = std::min<int64_t>(n, rg::kTcBitsV);
    const uint64_t dmin = o && o->heavy_degree > 0 ? (uint64_t)o->heavy_degree : (uint64_t)rg::kTcBitsDeg;
    unsigned long long* next_v = reinterpret_cast<unsigned long long*>(ctx->counts.as<int64_t>() + 21);
    unsigned long long* split = reinterpret_cast<unsigned long long*>(ctx->counts.as<int64_t>() + 22);
    RG_HIP(hipMemsetAsync(next_v, 0, sizeof(uint64_t), s));
    RG_TRY(rg::fill<unsigned long long>(ctx, split, 1, (unsigned long long)vcap));
    hipLaunchKernelGGL(rg::k_tc_split, rg::grid_for(ctx, vcap), 256, 0, s, g->off, vcap, dmin, split);
    hipLaunchKernelGGL(rg::k_tc_bits, ctx->num_cus * 2, rg::kTcBitsThreads, 0, s, g->off, g->nbr, split, next_v,
                       d_total);
    hipLaunchKernelGGL(rg::k_tc_flat, grid, 256, 0, s, g->off, g->nbr, n, d_total, split);
  }
  RG_HIP(hipGetLastError());
  RG_HIP(hipEventRecord(b, s));
  RG_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  RG_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  RG_HIP(hipMemcpy(total_out, d_total, sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (per_vertex) RG_HIP(hipMemcpy(per_vertex, tri.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
  if (ms_out) *ms_out = ms;
  return RG_OK;
}
