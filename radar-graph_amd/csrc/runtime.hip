// Context, device graph and buffer management of the radar_graph library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "device.h"

// Fixed device scratch of rg_bfs_noalloc (its frontier lists, byte map,
// heavy queue and counters): the loader places one copy on every GPU the
// library runs on, held until the process ends.  In host code the name is
// not a device address; the address comes from hipGetSymbolAddress, once
// per device (fixed_scratch_address), and only that address is used.
static __device__ __align__(256) unsigned char scratch[RG_FIXED_SCRATCH_BYTES];

namespace {
std::mutex g_fixed_mu;
constexpr int kMaxDevices = 64;
void* g_fixed_addr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
int fixed_scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_fixed_mu);
  if (device < 0 || device >= kMaxDevices)
    return rg::set_error(RG_EINVAL, "fixed scratch: device index above " + std::to_string(kMaxDevices - 1));
  if (!g_fixed_addr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return rg::set_error(RG_EDEVICE, std::string("fixed scratch: hipGetSymbolAddress failed: ") +
                                           hipGetErrorString(e));
    }
    g_fixed_addr[device] = p;
  }
  *out = g_fixed_addr[device];
  return RG_OK;
}
}  // namespace

namespace rg {

int DevBuf::ensure(size_t want) {
  if (want <= bytes && p) return RG_OK;
  if (p) {
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  size_t sz = std::max<size_t>(want, 256);
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return set_error(RG_ENOMEM, "device allocation of " + std::to_string(sz) + " bytes failed");
  }
  bytes = sz;
  return RG_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void hub_meta_release(rg_graph* g, HubMeta& h) {
  for (uint32_t* p : {h.bits, h.rank, h.ids, h.slot_of, h.filter})
    if (p && !borrowed(g, p)) (void)hipFree(p);
  h = HubMeta{};
}

void GraphCache::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  key.clear();
}

int ctx_frontiers(rg_ctx* ctx, int64_t n, Frontier& a, Frontier& b) {
  size_t cnt = (size_t)std::max<int64_t>(n, 64);
  RG_TRY(ctx->fr_ids[0].ensure(sizeof(uint32_t) * cnt));
  RG_TRY(ctx->fr_ids[1].ensure(sizeof(uint32_t) * cnt));
  if (ctx->fr_dense.bytes < cnt || !ctx->fr_dense.p) {
    RG_TRY(ctx->fr_dense.ensure(cnt));
    RG_HIP(hipMemsetAsync(ctx->fr_dense.p, 0, ctx->fr_dense.bytes, ctx->stream));
  }
  a = Frontier{};
  b = Frontier{};
  a.ids = ctx->fr_ids[0].as<uint32_t>();
  b.ids = ctx->fr_ids[1].as<uint32_t>();
  a.dense = b.dense = ctx->fr_dense.as<uint8_t>();
  return RG_OK;
}

hipEvent_t* timing_begin(rg_ctx* ctx) {
  if (!ctx->timing) return nullptr;
  if (ctx->ev_used == ctx->ev_pairs.size()) {
    rg_ctx::EvPair p;
    if (hipEventCreate(&p.e[0]) != hipSuccess || hipEventCreate(&p.e[1]) != hipSuccess) return nullptr;
    ctx->ev_pairs.push_back(p);
  }
  auto& p = ctx->ev_pairs[ctx->ev_used++];
  if (hipEventRecord(p.e[0], ctx->stream) != hipSuccess) return nullptr;
  return p.e;
}

int64_t ctx_device_bytes(const rg_ctx* ctx) {
  int64_t b = (int64_t)ctx->counts.bytes + (int64_t)ctx->fr_ids[0].bytes + (int64_t)ctx->fr_ids[1].bytes +
              (int64_t)ctx->fr_dense.bytes + (int64_t)ctx->bc_dense.bytes;
  for (const DevBuf& s : ctx->scratch) b += (int64_t)s.bytes;
  return b + ms_state_bytes(ctx->ms) + bcb_state_bytes(ctx->bcb);
}

void ctx_release_batch_state(rg_ctx* ctx, int which) {
  (void)hipStreamSynchronize(ctx->stream);
  if (which & 1) {
    for (auto& gcache : ctx->ms_graph) gcache.release();
    ms_state_free(ctx->ms);
    ctx->ms = nullptr;
  }
  if (which & 2) {
    bcb_state_free(ctx->bcb);
    ctx->bcb = nullptr;
  }
}

int ctx_copy(rg_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return RG_OK;
  RG_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
  RG_HIP(hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

int ctx_read_count(rg_ctx* ctx, const int64_t* dev, int64_t* host, int nwords) {
  RG_HIP(hipMemcpyAsync(ctx->h_counts, dev, sizeof(int64_t) * nwords, hipMemcpyDeviceToHost,
                        ctx->stream));
  RG_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(host, ctx->h_counts, sizeof(int64_t) * nwords);
  return RG_OK;
}

}  // namespace rg

using rg::set_error;

extern "C" {

const char* rg_version(void) { return "radar_graph 0.1 (gfx950)"; }

int rg_device_count(int* count) {
  if (!count) return set_error(RG_EINVAL, "rg_device_count: null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  *count = c;
  return RG_OK;
}

int rg_ctx_create(int device, rg_ctx** out) {
  if (!out) return set_error(RG_EINVAL, "rg_ctx_create: null out");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0) {
    (void)hipGetLastError();
    return set_error(RG_EDEVICE, "rg_ctx_create: no HIP device available");
  }
  if (device < 0 || device >= c) return set_error(RG_EINVAL, "rg_ctx_create: bad device index");
  RG_HIP(hipSetDevice(device));
  std::unique_ptr<rg_ctx> ctx(new rg_ctx);
  ctx->device = device;
  hipDeviceProp_t prop;
  RG_HIP(hipGetDeviceProperties(&prop, device));
  ctx->num_cus = prop.multiProcessorCount;
  ctx->max_lds_per_block = prop.sharedMemPerBlock;
  if (prop.maxSharedMemoryPerMultiProcessor > 0)
    ctx->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  RG_TRY(fixed_scratch_address(device, &ctx->fixed_scratch));
  RG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  RG_HIP(hipHostMalloc((void**)&ctx->h_counts, 64 * sizeof(int64_t), hipHostMallocDefault));
  RG_TRY(ctx->counts.ensure(64 * sizeof(int64_t)));
  *out = ctx.release();
  return RG_OK;
}

int rg_ctx_destroy(rg_ctx* ctx) {
  if (!ctx) return RG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->pr_graph.release();
  ctx->bfs_graph.release();
  for (auto& gcache : ctx->ms_graph) gcache.release();
  ctx->counts.release();
  for (auto& b : ctx->scratch) b.release();
  ctx->fr_ids[0].release();
  ctx->fr_ids[1].release();
  ctx->fr_dense.release();
  ctx->bc_dense.release();
  rg::ms_state_free(ctx->ms);
  ctx->ms = nullptr;
  rg::bcb_state_free(ctx->bcb);
  ctx->bcb = nullptr;
  for (auto& p : ctx->ev_pairs) {
    (void)hipEventDestroy(p.e[0]);
    (void)hipEventDestroy(p.e[1]);
  }
  if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RG_OK;
}

int rg_ctx_timing(rg_ctx* ctx, int enable) {
  if (!ctx) return set_error(RG_EINVAL, "rg_ctx_timing: null");
  RG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timing = enable != 0;
  ctx->ev_used = 0;
  return RG_OK;
}

int rg_ctx_graphs(rg_ctx* ctx, int enable) {
  if (!ctx) return set_error(RG_EINVAL, "rg_ctx_graphs: null");
  RG_HIP(hipStreamSynchronize(ctx->stream));
  ctx->graphs = (enable & RG_GRAPHS_LOOPS) != 0;
  ctx->bfs_device_levels = (enable & RG_GRAPHS_BFS_LEVELS) != 0;
  return RG_OK;
}

int rg_ctx_kernel_time(rg_ctx* ctx, int64_t* launches, double* total_ms) {
  if (!ctx) return set_error(RG_EINVAL, "rg_ctx_kernel_time: null");
  RG_HIP(hipStreamSynchronize(ctx->stream));
  double t = 0.0;
  for (size_t i = 0; i < ctx->ev_used; ++i) {
    float ms = 0.f;
    RG_HIP(hipEventElapsedTime(&ms, ctx->ev_pairs[i].e[0], ctx->ev_pairs[i].e[1]));
    t += ms;
  }
  if (launches) *launches = (int64_t)ctx->ev_used;
  if (total_ms) *total_ms = t;
  return RG_OK;
}

// Plain streaming copy, 16 bytes per lane per step: the reference point for
// HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int rg_device_copy_bandwidth(rg_ctx* ctx, int64_t bytes, int reps, double* gbps, double* ms_per_copy) {
  if (!ctx || bytes < 16 || reps < 1) return set_error(RG_EINVAL, "rg_device_copy_bandwidth: bad arguments");
  RG_HIP(hipSetDevice(ctx->device));
  const int64_t n16 = bytes / 16;
  rg::DevBuf a, b;
  RG_TRY(a.ensure((size_t)n16 * 16));
  RG_TRY(b.ensure((size_t)n16 * 16));
  RG_HIP(hipMemsetAsync(a.p, 1, (size_t)n16 * 16, ctx->stream));
  const int grid = ctx->num_cus * 8;
  hipLaunchKernelGGL(k_stream_copy, grid, 256, 0, ctx->stream, a.as<uint4>(), b.as<uint4>(), n16);
  hipEvent_t e0, e1;
  RG_HIP(hipEventCreate(&e0));
  RG_HIP(hipEventCreate(&e1));
  RG_HIP(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_stream_copy, grid, 256, 0, ctx->stream, a.as<uint4>(), b.as<uint4>(), n16);
  RG_HIP(hipGetLastError());
  RG_HIP(hipEventRecord(e1, ctx->stream));
  RG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  RG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double per = ms / reps;
  if (ms_per_copy) *ms_per_copy = per;
  if (gbps) *gbps = 2.0 * (double)(n16 * 16) / (per * 1e-3) / 1e9;   // read + write
  return RG_OK;
}

// Read-only stream: every 16-byte word of src is loaded once; a workgroup
// stores only when its XOR of everything it read equals an impossible
// pattern, so the loads cannot be dropped and (for data that is not that
// pattern) nothing is written.
__global__ void k_stream_read(const uint4* __restrict__ src, int64_t n16, uint4* __restrict__ sink) {
  uint4 x = make_uint4(0, 0, 0, 0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint4 v = src[i];
    x.x ^= v.x; x.y ^= v.y; x.z ^= v.z; x.w ^= v.w;
  }
  if (x.x == 0xdeadbeefu && x.y == 0x01234567u && x.z == 0x89abcdefu && x.w == 0x0badf00du) sink[blockIdx.x] = x;
}

// Random 8-byte gathers: count loads of src[hash(i) mod n8], the hash a
// splitmix64 finalizer, so each load lands on its own line of a buffer far
// larger than the caches; the same XOR guard keeps stores away.
__global__ void k_gather_read(const uint64_t* __restrict__ src, int64_t n8, int64_t count, uint64_t* __restrict__ sink) {
  uint64_t x = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t h = (uint64_t)i + 0x9e3779b97f4a7c15ull;
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
    h ^= h >> 31;
    x ^= src[h % (uint64_t)n8];
  }
  if (x == 0xdeadbeef01234567ull) sink[blockIdx.x] = x;
}

int rg_device_read_bandwidth(rg_ctx* ctx, int64_t bytes, int reps, int64_t gathers, double* gbps,
                             double* ms_per_read, double* ms_per_gather_pass) {
  if (!ctx || bytes < 16 || reps < 1 || gathers < 0)
    return set_error(RG_EINVAL, "rg_device_read_bandwidth: bad arguments");
  RG_HIP(hipSetDevice(ctx->device));
  const int64_t n16 = bytes / 16;
  rg::DevBuf a, sink;
  const int grid = ctx->num_cus * 8;
  RG_TRY(a.ensure((size_t)n16 * 16));
  RG_TRY(sink.ensure(sizeof(uint4) * grid));
  RG_HIP(hipMemsetAsync(a.p, 1, (size_t)n16 * 16, ctx->stream));
  hipEvent_t e0, e1, e2;
  RG_HIP(hipEventCreate(&e0));
  RG_HIP(hipEventCreate(&e1));
  RG_HIP(hipEventCreate(&e2));
  hipLaunchKernelGGL(k_stream_read, grid, 256, 0, ctx->stream, a.as<uint4>(), n16, sink.as<uint4>());
  RG_HIP(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_stream_read, grid, 256, 0, ctx->stream, a.as<uint4>(), n16, sink.as<uint4>());
  RG_HIP(hipEventRecord(e1, ctx->stream));
  for (int r = 0; r < (gathers ? reps : 0); ++r)
    hipLaunchKernelGGL(k_gather_read, grid, 256, 0, ctx->stream, a.as<uint64_t>(), n16 * 2, gathers,
                       sink.as<uint64_t>());
  RG_HIP(hipEventRecord(e2, ctx->stream));
  RG_HIP(hipGetLastError());
  RG_HIP(hipEventSynchronize(e2));
  float ms = 0.f, msg = 0.f;
  RG_HIP(hipEventElapsedTime(&ms, e0, e1));
  RG_HIP(hipEventElapsedTime(&msg, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  const double per = ms / reps;
  if (ms_per_read) *ms_per_read = per;
  if (gbps) *gbps = (double)(n16 * 16) / (per * 1e-3) / 1e9;
  if (ms_per_gather_pass) *ms_per_gather_pass = gathers ? msg / reps : 0.0;
  return RG_OK;
}

int rg_ctx_sync(rg_ctx* ctx) {
  if (!ctx) return set_error(RG_EINVAL, "rg_ctx_sync: null");
  RG_HIP(hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

static int alloc_copy(void** dst, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dst = nullptr;
    return set_error(RG_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  if (bytes) RG_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s));
  return RG_OK;
}

int rg_graph_upload(rg_ctx* ctx, const rg_hgraph* hg, rg_graph** out) {
  if (!ctx || !hg || !out) return set_error(RG_EINVAL, "rg_graph_upload: null argument");
  RG_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<rg_graph> g(new rg_graph);
  g->ctx = ctx;
  g->n = hg->n;
  g->m = hg->m;
  g->symmetric = hg->symmetric;
  int rc = alloc_copy((void**)&g->off, hg->off.data(), sizeof(uint64_t) * (hg->n + 1), ctx->stream);
  if (!rc) rc = alloc_copy((void**)&g->nbr, hg->nbr.data(), sizeof(uint32_t) * hg->m, ctx->stream);
  if (!rc && !hg->symmetric) {
    rc = alloc_copy((void**)&g->in_off, hg->in_off.data(), sizeof(uint64_t) * (hg->n + 1), ctx->stream);
    if (!rc) rc = alloc_copy((void**)&g->in_nbr, hg->in_nbr.data(), sizeof(uint32_t) * hg->m, ctx->stream);
  } else if (!rc) {
    g->in_off = g->off;
    g->in_nbr = g->nbr;
  }
  if (rc) {
    rg_graph_free(g.release());
    return rc;
  }
  RG_HIP(hipStreamSynchronize(ctx->stream));
  *out = g.release();
  return RG_OK;
}

int rg_graph_info(const rg_graph* g, int64_t* n, int64_t* m, int* symmetric) {
  if (!g) return set_error(RG_EINVAL, "rg_graph_info: null graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (symmetric) *symmetric = g->symmetric ? 1 : 0;
  return RG_OK;
}

int rg_graph_download(const rg_graph* g, int in_edges, uint64_t* offsets, uint32_t* nbrs) {
  if (!g || !offsets || (g->m > 0 && !nbrs)) return set_error(RG_EINVAL, "rg_graph_download: null argument");
  RG_HIP(hipSetDevice(g->ctx->device));
  const uint64_t* o = in_edges ? g->in_off : g->off;
  const uint32_t* b = in_edges ? g->in_nbr : g->nbr;
  RG_HIP(hipMemcpyAsync(offsets, o, sizeof(uint64_t) * (g->n + 1), hipMemcpyDeviceToHost, g->ctx->stream));
  if (g->m) RG_HIP(hipMemcpyAsync(nbrs, b, sizeof(uint32_t) * g->m, hipMemcpyDeviceToHost, g->ctx->stream));
  RG_HIP(hipStreamSynchronize(g->ctx->stream));
  return RG_OK;
}

int rg_graph_transpose(rg_graph* g) {
  if (!g) return set_error(RG_EINVAL, "rg_graph_transpose: null graph");
  if (g->symmetric) return RG_OK;
  std::swap(g->off, g->in_off);
  std::swap(g->nbr, g->in_nbr);
  g->transposed = !g->transposed;
  return RG_OK;
}

int rg_graph_free(rg_graph* g) {
  if (!g) return RG_OK;
  if (g->ctx) (void)hipSetDevice(g->ctx->device);
  // A view frees only what it built itself (not borrowed from its base);
  // the CSR and CSC are always the base's.
  if (!g->view) {
    if (g->in_off && g->in_off != g->off) (void)hipFree(g->in_off);
    if (g->in_nbr && g->in_nbr != g->nbr) (void)hipFree(g->in_nbr);
    if (g->off) (void)hipFree(g->off);
    if (g->nbr) (void)hipFree(g->nbr);
  }
  for (rg::HubMeta& h : g->hubs) rg::hub_meta_release(g, h);
  if (g->first_in && !rg::borrowed(g, g->first_in)) (void)hipFree(g->first_in);
  if (g->heavy_items && !rg::borrowed(g, g->heavy_items)) (void)hipFree(g->heavy_items);
  for (void* p : {g->blocked, g->blocked_ms, g->blocked_ms_t})
    if (p && !rg::borrowed(g, p)) rg::blocked_free(p);
  delete g;
  return RG_OK;
}

}  // extern "C"
