// DegreeSorting on the GPU: the reference's preprocessGraph
// (push-new-orig Ligra/ligra/IO.h:729-944; relaxed copy in
// degree-sorting/Ligra/ligra/IO.h:729-894) and GAP's RelabelByDegree
// (push-new-orig GAP/src/builder.h:246-282).
//
//   1. key[v] = degree(v) << 32 | v, radix-sorted descending: the order of
//      std::greater on (degree, id) pairs -- degree descending, ties by id
//      descending -- with unique keys, so no stability question arises;
//   2. new_ids[order[r]] = r;
//   3. new offsets = exclusive scan of the degrees in the new order;
//   4. every list is copied to its new place with neighbours relabelled,
//      keeping the original order (relaxed) or sorted afterwards (strict).
// Steps 1, 3 and the optional segmented sort use rocPRIM's radix sort / scan
// through hipCUB; the relabel gather is a hand-written kernel.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <memory>
#include <vector>

#include "device.h"

namespace rg {
namespace {

__global__ void k_deg_keys(const uint64_t* off, int64_t n, uint64_t* keys) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    keys[v] = ((off[v + 1] - off[v]) << 32) | (uint64_t)v;
}

__global__ void k_new_ids(const uint64_t* sorted_keys, int64_t n, uint32_t* new_ids, uint32_t* inv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = (uint32_t)(sorted_keys[r] & 0xffffffffu);
    new_ids[v] = (uint32_t)r;
    inv[r] = v;
  }
}

__global__ void k_new_degrees(const uint64_t* off, const uint32_t* inv, int64_t n, uint64_t* ndeg) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = inv[r];
    ndeg[r] = off[v + 1] - off[v];
  }
}

// One wavefront per new vertex: lanes copy the old list (coalesced) and map
// every neighbour through new_ids.
__global__ void k_relabel_lists(const uint64_t* off, const uint32_t* nbr, const uint32_t* inv,
                                const uint32_t* new_ids, const uint64_t* noff, int64_t n,
                                uint32_t* nnbr) {
  const int lane = (int)__lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t r = wid; r < n; r += nw) {
    uint32_t v = inv[r];
    uint64_t b = off[v], e = off[v + 1], o = noff[r];
    for (uint64_t j = b + lane; j < e; j += kWave) nnbr[o + (j - b)] = new_ids[nbr[j]];
  }
}

__global__ void k_set_last(uint64_t* noff, int64_t n, uint64_t m) { noff[n] = m; }

// Offsets of vertices [v0, v0 + cnt] relative to the first one (the
// segmented sort takes int item counts: lists are sorted in chunks of fewer
// than 2^31 arcs).
__global__ void k_rel_offsets(const uint64_t* noff, int64_t v0, int64_t cnt, uint64_t* rel) {
  const uint64_t base = noff[v0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= cnt; i += (int64_t)gridDim.x * blockDim.x)
    rel[i] = noff[v0 + i] - base;
}

int grid(rg_ctx* ctx, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 16));
}

}  // namespace

// Sorts every list of (noff, nbr) into sorted (the segmented sort takes int
// item and segment counts: graphs of 2^31 arcs or more, or any graph when
// chunk_arcs > 0, are sorted in chunks of consecutive vertices holding at
// most chunk_arcs arcs each, default 2^30, with offsets relative to the
// chunk; one list longer than that is a chunk of its own).
int sort_lists(rg_ctx* ctx, const uint64_t* noff, const uint32_t* nbr, int64_t n, int64_t m,
               uint64_t chunk_arcs, DevBuf& tmp, uint32_t* sorted) {
  hipStream_t s = ctx->stream;
  if (m <= (int64_t)INT32_MAX && chunk_arcs == 0) {
    size_t sb = 0;
    RG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sb, nbr, sorted, (int)m, (int)n, noff, noff + 1,
                                                      0, 32, s));
    RG_TRY(tmp.ensure(sb));
    RG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, sb, nbr, sorted, (int)m, (int)n, noff, noff + 1, 0,
                                                      32, s));
    return RG_OK;
  }
  std::vector<uint64_t> h(n + 1);
  RG_HIP(hipMemcpyAsync(h.data(), noff, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, s));
  RG_HIP(hipStreamSynchronize(s));
  DevBuf rel;
  const uint64_t chunk = chunk_arcs ? chunk_arcs : (uint64_t)1 << 30;
  for (int64_t v0 = 0; v0 < n;) {
    int64_t v1 = v0 + 1;
    while (v1 < n && h[v1 + 1] - h[v0] <= chunk) ++v1;
    const uint64_t items = h[v1] - h[v0];
    if (items > (uint64_t)INT32_MAX) return set_error(RG_EUNSUPPORTED, "degree sort: a list of 2^31 arcs");
    if (items > 1) {
      const int64_t cnt = v1 - v0;
      RG_TRY(rel.ensure(sizeof(uint64_t) * (cnt + 1)));
      hipLaunchKernelGGL(k_rel_offsets, grid(ctx, cnt + 1), 256, 0, s, noff, v0, cnt, rel.as<uint64_t>());
      const uint64_t* rb = rel.as<uint64_t>();
      size_t sb = 0;
      RG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sb, nbr + h[v0], sorted + h[v0], (int)items,
                                                        (int)cnt, rb, rb + 1, 0, 32, s));
      RG_TRY(tmp.ensure(sb));
      RG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, sb, nbr + h[v0], sorted + h[v0], (int)items,
                                                        (int)cnt, rb, rb + 1, 0, 32, s));
      // rel is rewritten by the next chunk: this chunk's sort must be done.
      RG_HIP(hipStreamSynchronize(s));
    } else if (items == 1) {
      RG_HIP(hipMemcpyAsync(sorted + h[v0], nbr + h[v0], sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    v0 = v1;
  }
  return RG_OK;
}

namespace {

// Builds one relabelled CSR (offsets + lists) from (off, nbr); on success
// the caller owns *noff_out and *nnbr_out.
int relabel_csr(rg_ctx* ctx, const uint64_t* off, const uint32_t* nbr, int64_t n, int64_t m,
                const uint32_t* inv, const uint32_t* new_ids, bool strict, uint64_t chunk_arcs,
                uint64_t** noff_out, uint32_t** nnbr_out) {
  hipStream_t s = ctx->stream;
  DevBuf noff, nnbr, ndeg, tmp;
  RG_TRY(noff.ensure(sizeof(uint64_t) * (n + 1)));
  RG_TRY(nnbr.ensure(sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  RG_TRY(ndeg.ensure(sizeof(uint64_t) * n));
  hipLaunchKernelGGL(k_new_degrees, grid(ctx, n), 256, 0, s, off, inv, n, ndeg.as<uint64_t>());
  size_t tb = 0;
  RG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ndeg.as<uint64_t>(), noff.as<uint64_t>(), (int)n, s));
  RG_TRY(tmp.ensure(tb));
  RG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, ndeg.as<uint64_t>(), noff.as<uint64_t>(), (int)n, s));
  hipLaunchKernelGGL(k_set_last, 1, 1, 0, s, noff.as<uint64_t>(), n, (uint64_t)m);
  hipLaunchKernelGGL(k_relabel_lists, grid(ctx, n), 256, 0, s, off, nbr, inv, new_ids, noff.as<uint64_t>(), n,
                     nnbr.as<uint32_t>());
  RG_HIP(hipGetLastError());
  if (strict && m > 0) {
    DevBuf sorted;
    RG_TRY(sorted.ensure(sizeof(uint32_t) * m));
    RG_TRY(sort_lists(ctx, noff.as<uint64_t>(), nnbr.as<uint32_t>(), n, m, chunk_arcs, tmp, sorted.as<uint32_t>()));
    RG_HIP(hipStreamSynchronize(s));
    std::swap(nnbr.p, sorted.p);
    std::swap(nnbr.bytes, sorted.bytes);
  }
  RG_HIP(hipStreamSynchronize(s));
  *noff_out = noff.as<uint64_t>();
  *nnbr_out = nnbr.as<uint32_t>();
  noff.p = nullptr;
  nnbr.p = nullptr;
  return RG_OK;
}

}  // namespace
}  // namespace rg

using rg::set_error;

extern "C" int rg_graph_degree_sort(rg_graph* g, int by_outdeg, int sort_neighbors, rg_graph** out,
                                    uint32_t* new_ids_host) {
  if (!g || !out) return set_error(RG_EINVAL, "rg_graph_degree_sort: null argument");
  if (g->n > (int64_t)INT32_MAX) return set_error(RG_EUNSUPPORTED, "rg_graph_degree_sort: 2^31 or more vertices");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = g->n, m = g->m;
  rg::DevBuf keys, skeys, tmp, ids, inv;
  RG_TRY(keys.ensure(sizeof(uint64_t) * n));
  RG_TRY(skeys.ensure(sizeof(uint64_t) * n));
  RG_TRY(ids.ensure(sizeof(uint32_t) * n));
  RG_TRY(inv.ensure(sizeof(uint32_t) * n));
  // -preprocess 1 sorts by in-degree, 0 by out-degree (ligra.h:568-569);
  // symmetric graphs have one degree.
  const uint64_t* key_off = (by_outdeg || g->symmetric) ? g->off : g->in_off;
  hipLaunchKernelGGL(rg::k_deg_keys, rg::grid(ctx, n), 256, 0, s, key_off, n, keys.as<uint64_t>());
  size_t tb = 0;
  RG_HIP(hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tb, keys.as<uint64_t>(), skeys.as<uint64_t>(),
                                                     (int)n, 0, 64, s));
  RG_TRY(tmp.ensure(tb));
  RG_HIP(hipcub::DeviceRadixSort::SortKeysDescending(tmp.p, tb, keys.as<uint64_t>(), skeys.as<uint64_t>(),
                                                     (int)n, 0, 64, s));
  hipLaunchKernelGGL(rg::k_new_ids, rg::grid(ctx, n), 256, 0, s, skeys.as<uint64_t>(), n,
                     ids.as<uint32_t>(), inv.as<uint32_t>());
  RG_HIP(hipGetLastError());
  keys.release();
  skeys.release();
  tmp.release();
  std::unique_ptr<rg_graph> ng(new rg_graph);
  ng->ctx = ctx;
  ng->n = n;
  ng->m = m;
  ng->symmetric = g->symmetric;
  ng->degree_sorted = true;
  // RG_DEGREE_SORT_CHUNK (arcs) forces the chunked list sort that graphs of
  // 2^31 arcs or more take.
  const char* chunk_env = std::getenv("RG_DEGREE_SORT_CHUNK");
  const uint64_t chunk_arcs = chunk_env ? std::strtoull(chunk_env, nullptr, 10) : 0;
  int rc = rg::relabel_csr(ctx, g->off, g->nbr, n, m, inv.as<uint32_t>(), ids.as<uint32_t>(),
                           sort_neighbors != 0, chunk_arcs, &ng->off, &ng->nbr);
  if (rc == RG_OK) {
    if (g->symmetric) {
      ng->in_off = ng->off;
      ng->in_nbr = ng->nbr;
    } else {
      rc = rg::relabel_csr(ctx, g->in_off, g->in_nbr, n, m, inv.as<uint32_t>(), ids.as<uint32_t>(),
                           sort_neighbors != 0, chunk_arcs, &ng->in_off, &ng->in_nbr);
    }
  }
  if (rc == RG_OK && new_ids_host)
    rc = hipMemcpy(new_ids_host, ids.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost) == hipSuccess
             ? RG_OK
             : set_error(RG_EDEVICE, "rg_graph_degree_sort: copy of new_ids failed");
  ids.release();
  inv.release();
  if (rc != RG_OK) {
    rg_graph_free(ng.release());
    return rc;
  }
  *out = ng.release();
  return RG_OK;
}
