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
// Mapping: one wavefront per vertex u; its lanes take the arcs (u, v) of
// N+(u) and each merges N+(u) with N+(v).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"

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

}  // namespace
}  // namespace rg

using rg::set_error;

extern "C" int rg_tc(rg_graph* g, const rg_opts* o, uint64_t* total_out, uint64_t* per_vertex,
                     double* ms_out) {
  (void)o;
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
  hipLaunchKernelGGL(rg::k_tc, grid, 256, 0, s, g->off, g->nbr, n, d_total,
                     per_vertex ? tri.as<unsigned long long>() : nullptr);
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
