// Fast ingest: graph files straight into device memory, and device graphs
// back to files.
//
// The reference reads a binary graph into host arrays (Ligra's binary
// reader, push-new-orig Ligra/ligra/IO.h:477-626; GAP's serialized-graph
// reader, GAP/src/reader.h:251-304) and its apps then run on those host
// arrays.  Here the file is memory-mapped and streamed to the device through
// two pinned staging buffers: worker threads copy the next chunk from the
// page cache into one buffer while the DMA engine moves the other, so no host
// CSR is ever built.  The writer (GAP/src/writer.h:39-69 for .sg) streams the
// other way.  A graph saved after DegreeSorting keeps its new_ids in a
// sidecar file (<path>.newids, n little-endian uint32 values: new_ids[orig]),
// so a serving or benchmark restart loads the relabelled graph instead of
// generating, building and sorting it again.
//
// Device layout as rg_graph_upload makes it: uint64 offsets[n+1], uint32
// neighbours[m], and for directed graphs the in-edge CSR (read from a
// directed .sg file; built on the device for directed Ligra binary files,
// whose in-lists come out sorted by source as the host reader makes them).
// Every loaded CSR is checked on the device before use (offsets start at 0,
// never decrease and end at m, neighbours < n), since the traversal kernels
// index with these values.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "device.h"

namespace rg {
namespace {

constexpr size_t kStageBytes = size_t(64) << 20;   // per pinned staging buffer

// A read-only memory-mapped file.
struct Mapped {
  int fd = -1;
  const char* p = nullptr;
  size_t len = 0;
  ~Mapped() {
    if (p && len) munmap(const_cast<char*>(p), len);
    if (fd >= 0) close(fd);
  }
  int open_file(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return set_error(RG_EIO, "cannot read " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) return set_error(RG_EIO, "cannot stat " + path);
    len = (size_t)st.st_size;
    if (len == 0) return RG_OK;
    void* m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) return set_error(RG_EIO, "cannot map " + path);
    p = static_cast<const char*>(m);
    madvise(m, len, MADV_SEQUENTIAL);
    return RG_OK;
  }
};

// Two pinned staging buffers and their completion events on the context
// stream.
struct Stager {
  rg_ctx* ctx;
  char* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  explicit Stager(rg_ctx* c) : ctx(c) {}
  ~Stager() {
    (void)hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < 2; ++k) {
      if (pin[k]) (void)hipHostFree(pin[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
  }
  int init() {
    for (int k = 0; k < 2; ++k) {
      if (hipHostMalloc((void**)&pin[k], kStageBytes, hipHostMallocDefault) != hipSuccess) {
        pin[k] = nullptr;
        (void)hipGetLastError();
        return set_error(RG_ENOMEM, "pinned staging allocation failed");
      }
      RG_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    return RG_OK;
  }
  // Host bytes -> device, chunk by chunk through the two buffers.
  int to_device(void* dev, const char* src, size_t bytes) {
    char* d = static_cast<char*>(dev);
    int k = 0;
    for (size_t pos = 0; pos < bytes; pos += kStageBytes, k ^= 1) {
      const size_t len = std::min(kStageBytes, bytes - pos);
      if (used[k]) RG_HIP(hipEventSynchronize(ev[k]));   // the DMA out of this buffer is done
      char* buf = pin[k];
      const char* s = src + pos;
      parallel_chunks(0, (int64_t)len, [&](int64_t b, int64_t e) { std::memcpy(buf + b, s + b, (size_t)(e - b)); });
      RG_HIP(hipMemcpyAsync(d + pos, buf, len, hipMemcpyHostToDevice, ctx->stream));
      RG_HIP(hipEventRecord(ev[k], ctx->stream));
      used[k] = true;
    }
    return RG_OK;
  }
  // Device bytes -> file: chunk i is copied out while chunk i - 1 is written.
  int to_file(FILE* f, const void* dev, size_t bytes) {
    const char* d = static_cast<const char*>(dev);
    size_t prev_len = 0;
    int k = 0;
    for (size_t pos = 0;; pos += kStageBytes, k ^= 1) {
      const size_t len = pos < bytes ? std::min(kStageBytes, bytes - pos) : 0;
      if (len) {
        RG_HIP(hipMemcpyAsync(pin[k], d + pos, len, hipMemcpyDeviceToHost, ctx->stream));
        RG_HIP(hipEventRecord(ev[k], ctx->stream));
      }
      if (prev_len) {
        RG_HIP(hipEventSynchronize(ev[k ^ 1]));
        if (std::fwrite(pin[k ^ 1], 1, prev_len, f) != prev_len) return set_error(RG_EIO, "short write");
      }
      if (!len) break;
      prev_len = len;
    }
    return RG_OK;
  }
};

__global__ void k_widen_offsets(const uint32_t* o32, int64_t n, uint64_t m, uint64_t* o64) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x)
    o64[i] = i < n ? (uint64_t)o32[i] : m;
}

// Flags (bits of *bad): 1 offsets[0] != 0, 2 offsets[n] != m, 4 a decreasing
// offset, 8 a neighbour id >= n.
__global__ void k_check_csr(const uint64_t* off, const uint32_t* nbr, int64_t n, uint64_t m, unsigned int* bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  unsigned int f = 0;
  if (t == 0) {
    if (off[0] != 0) f |= 1;
    if (off[n] != m) f |= 2;
  }
  for (int64_t v = t; v < n; v += stride)
    if (off[v + 1] < off[v]) f |= 4;
  for (uint64_t j = (uint64_t)t; j < m; j += (uint64_t)stride)
    if (nbr[j] >= (uint64_t)n) f |= 8;
  if (f) atomicOr(bad, f);
}

// In-degree count, then the scatter of one source list per wavefront
// (positions by atomics; the lists are sorted afterwards).
__global__ void k_in_count(const uint32_t* nbr, uint64_t m, uint64_t* cnt) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[nbr[j]]), 1ull);
}

__global__ void k_in_scatter(const uint64_t* off, const uint32_t* nbr, int64_t n, const uint64_t* in_off,
                             uint64_t* fill, uint32_t* in_nbr) {
  const int lane = (int)__lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t u = wid; u < n; u += nw)
    for (uint64_t j = off[u] + lane; j < off[u + 1]; j += kWave) {
      const uint32_t d = nbr[j];
      const uint64_t pos = atomicAdd(reinterpret_cast<unsigned long long*>(&fill[d]), 1ull);
      in_nbr[in_off[d] + pos] = (uint32_t)u;
    }
}

int grid_for(rg_ctx* ctx, int64_t items) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, (int64_t)ctx->num_cus * 16));
}

int check_csr(rg_ctx* ctx, const uint64_t* off, const uint32_t* nbr, int64_t n, int64_t m, const std::string& what) {
  unsigned int* bad = reinterpret_cast<unsigned int*>(ctx->counts.as<int64_t>() + 60);
  RG_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream));
  hipLaunchKernelGGL(k_check_csr, grid_for(ctx, std::max<int64_t>(n, m)), 256, 0, ctx->stream, off, nbr, n,
                     (uint64_t)m, bad);
  RG_HIP(hipGetLastError());
  unsigned int h = 0;
  RG_TRY(ctx_copy(ctx, &h, bad, sizeof(h), hipMemcpyDeviceToHost));
  if (h) return set_error(RG_EFORMAT, "Bad input file: " + what + " is not a valid CSR (check flags " +
                                          std::to_string(h) + ")");
  return RG_OK;
}

// Device CSC of the CSR (off, nbr): in-lists sorted by source.
int build_in_edges(rg_ctx* ctx, const uint64_t* off, const uint32_t* nbr, int64_t n, int64_t m, DevBuf& in_off,
                   DevBuf& in_nbr) {
  hipStream_t s = ctx->stream;
  DevBuf cnt, tmp, unsorted;
  RG_TRY(in_off.ensure(sizeof(uint64_t) * (n + 1)));
  RG_TRY(in_nbr.ensure(sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  RG_TRY(cnt.ensure(sizeof(uint64_t) * (n + 1)));
  RG_TRY(unsorted.ensure(sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  RG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(uint64_t) * (n + 1), s));
  hipLaunchKernelGGL(k_in_count, grid_for(ctx, m), 256, 0, s, nbr, (uint64_t)m, cnt.as<uint64_t>());
  size_t tb = 0;
  RG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<uint64_t>(), in_off.as<uint64_t>(), (int)(n + 1), s));
  RG_TRY(tmp.ensure(tb));
  RG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.as<uint64_t>(), in_off.as<uint64_t>(), (int)(n + 1), s));
  RG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(uint64_t) * (n + 1), s));
  hipLaunchKernelGGL(k_in_scatter, grid_for(ctx, n * kWave), 256, 0, s, off, nbr, n, in_off.as<uint64_t>(),
                     cnt.as<uint64_t>(), unsorted.as<uint32_t>());
  RG_HIP(hipGetLastError());
  cnt.release();
  if (m > 0) RG_TRY(sort_lists(ctx, in_off.as<uint64_t>(), unsorted.as<uint32_t>(), n, m, 0, tmp, in_nbr.as<uint32_t>()));
  RG_HIP(hipStreamSynchronize(s));
  return RG_OK;
}

struct FileShape {
  int64_t n = 0, m = 0;
  bool directed = false;
  size_t hdr = 0;   // .sg: bytes before the out offsets
};

int sg_shape(const Mapped& f, const std::string& path, FileShape& sh) {
  // bool directed, int64 edges, int64 nodes (GAP writer.h:45-50)
  sh.hdr = 1 + 8 + 8;
  if (f.len < sh.hdr) return set_error(RG_EFORMAT, "Bad input file: " + path);
  sh.directed = f.p[0] != 0;
  std::memcpy(&sh.m, f.p + 1, 8);
  std::memcpy(&sh.n, f.p + 9, 8);
  if (sh.n <= 0 || sh.n >= (int64_t)INT32_MAX || sh.m < 0) return set_error(RG_EFORMAT, "Bad input file: " + path);
  const size_t part = (size_t)(sh.n + 1) * 8 + (size_t)sh.m * 4;
  if (f.len < sh.hdr + part * (sh.directed ? 2 : 1)) return set_error(RG_EFORMAT, "Bad input file: " + path);
  return RG_OK;
}

int ligra_bin_shape(const std::string& prefix, Mapped& cfg, Mapped& adj, Mapped& idx, FileShape& sh) {
  RG_TRY(cfg.open_file(prefix + ".config"));
  RG_TRY(adj.open_file(prefix + ".adj"));
  RG_TRY(idx.open_file(prefix + ".idx"));
  const std::string c(cfg.p ? cfg.p : "", cfg.len);
  sh.n = std::atoll(c.c_str());
  if (sh.n <= 0 || sh.n >= (int64_t)INT32_MAX) return set_error(RG_EFORMAT, "Bad input file");
  if (idx.len != (size_t)sh.n * sizeof(uint32_t)) return set_error(RG_EFORMAT, "File size wrong");
  sh.m = (int64_t)(adj.len / sizeof(uint32_t));
  return RG_OK;
}

int read_new_ids(const std::string& path, int64_t n, uint32_t* out, bool* present) {
  *present = false;
  Mapped f;
  if (access((path + ".newids").c_str(), R_OK) != 0) return RG_OK;
  RG_TRY(f.open_file(path + ".newids"));
  if (f.len != (size_t)n * sizeof(uint32_t)) return set_error(RG_EFORMAT, "new_ids file size wrong: " + path);
  *present = true;
  if (out) std::memcpy(out, f.p, f.len);
  return RG_OK;
}

}  // namespace
}  // namespace rg

using rg::set_error;

extern "C" {

int rg_graph_file_info(const char* path, int format, int64_t* n, int64_t* m, int* symmetric, int* has_new_ids) {
  if (!path) return set_error(RG_EINVAL, "rg_graph_file_info: null path");
  rg::FileShape sh;
  const std::string p(path);
  int sym = 1;
  if (format == RG_FMT_SG) {
    rg::Mapped f;
    RG_TRY(f.open_file(p));
    RG_TRY(rg::sg_shape(f, p, sh));
    sym = sh.directed ? 0 : 1;
  } else if (format == RG_FMT_ADJ_BINARY) {
    rg::Mapped cfg, adj, idx;
    RG_TRY(rg::ligra_bin_shape(p, cfg, adj, idx, sh));
    sym = -1;   // the file does not say (the -s flag does)
  } else {
    return set_error(RG_EUNSUPPORTED, "rg_graph_file_info: binary formats only (.sg, Ligra .adj/.idx)");
  }
  bool ids = false;
  RG_TRY(rg::read_new_ids(p, sh.n, nullptr, &ids));
  if (n) *n = sh.n;
  if (m) *m = sh.m;
  if (symmetric) *symmetric = sym;
  if (has_new_ids) *has_new_ids = ids ? 1 : 0;
  return RG_OK;
}

int rg_graph_load(rg_ctx* ctx, const char* path, int format, int symmetric, uint32_t* new_ids, rg_graph** out) {
  if (!ctx || !path || !out) return set_error(RG_EINVAL, "rg_graph_load: null argument");
  RG_HIP(hipSetDevice(ctx->device));
  const std::string p(path);
  rg::Stager st(ctx);
  RG_TRY(st.init());
  std::unique_ptr<rg_graph, int (*)(rg_graph*)> g(new rg_graph, rg_graph_free);
  g->ctx = ctx;
  rg::FileShape sh;
  rg::DevBuf off, nbr, in_off, in_nbr;
  if (format == RG_FMT_SG) {
    rg::Mapped f;
    RG_TRY(f.open_file(p));
    RG_TRY(rg::sg_shape(f, p, sh));
    const size_t ob = sizeof(uint64_t) * (sh.n + 1), nb = sizeof(uint32_t) * sh.m;
    RG_TRY(off.ensure(ob));
    RG_TRY(nbr.ensure(std::max<size_t>(nb, 4)));
    RG_TRY(st.to_device(off.p, f.p + sh.hdr, ob));
    RG_TRY(st.to_device(nbr.p, f.p + sh.hdr + ob, nb));
    if (sh.directed) {
      RG_TRY(in_off.ensure(ob));
      RG_TRY(in_nbr.ensure(std::max<size_t>(nb, 4)));
      RG_TRY(st.to_device(in_off.p, f.p + sh.hdr + ob + nb, ob));
      RG_TRY(st.to_device(in_nbr.p, f.p + sh.hdr + 2 * ob + nb, nb));
    }
    RG_HIP(hipStreamSynchronize(ctx->stream));
  } else if (format == RG_FMT_ADJ_BINARY) {
    rg::Mapped cfg, adj, idx;
    RG_TRY(rg::ligra_bin_shape(p, cfg, adj, idx, sh));
    sh.directed = !symmetric;
    rg::DevBuf o32;
    RG_TRY(o32.ensure(sizeof(uint32_t) * sh.n));
    RG_TRY(off.ensure(sizeof(uint64_t) * (sh.n + 1)));
    RG_TRY(nbr.ensure(std::max<size_t>(sizeof(uint32_t) * sh.m, 4)));
    RG_TRY(st.to_device(o32.p, idx.p, sizeof(uint32_t) * sh.n));
    RG_TRY(st.to_device(nbr.p, adj.p, sizeof(uint32_t) * sh.m));
    hipLaunchKernelGGL(rg::k_widen_offsets, rg::grid_for(ctx, sh.n + 1), 256, 0, ctx->stream, o32.as<uint32_t>(),
                       sh.n, (uint64_t)sh.m, off.as<uint64_t>());
    RG_HIP(hipGetLastError());
    RG_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    return set_error(RG_EUNSUPPORTED, "rg_graph_load: binary formats only (.sg, Ligra .adj/.idx)");
  }
  RG_TRY(rg::check_csr(ctx, off.as<uint64_t>(), nbr.as<uint32_t>(), sh.n, sh.m, p));
  if (sh.directed && format == RG_FMT_SG)
    RG_TRY(rg::check_csr(ctx, in_off.as<uint64_t>(), in_nbr.as<uint32_t>(), sh.n, sh.m, p + " (in-edges)"));
  else if (sh.directed)
    RG_TRY(rg::build_in_edges(ctx, off.as<uint64_t>(), nbr.as<uint32_t>(), sh.n, sh.m, in_off, in_nbr));
  bool ids = false;
  RG_TRY(rg::read_new_ids(p, sh.n, new_ids, &ids));
  g->n = sh.n;
  g->m = sh.m;
  g->symmetric = !sh.directed;
  g->degree_sorted = ids;
  g->off = off.as<uint64_t>();
  g->nbr = nbr.as<uint32_t>();
  off.p = nullptr;
  nbr.p = nullptr;
  if (sh.directed) {
    g->in_off = in_off.as<uint64_t>();
    g->in_nbr = in_nbr.as<uint32_t>();
    in_off.p = nullptr;
    in_nbr.p = nullptr;
  } else {
    g->in_off = g->off;
    g->in_nbr = g->nbr;
  }
  *out = g.release();
  return RG_OK;
}

int rg_graph_save(const rg_graph* g, const char* path, int format, const uint32_t* new_ids) {
  if (!g || !path) return set_error(RG_EINVAL, "rg_graph_save: null argument");
  if (format != RG_FMT_SG && format != RG_FMT_ADJ_BINARY)
    return set_error(RG_EUNSUPPORTED, "rg_graph_save: binary formats only (.sg, Ligra .adj/.idx)");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  const std::string p(path);
  rg::Stager st(ctx);
  RG_TRY(st.init());
  // The stored orientation: the graph's original one (a transposed
  // directed graph is saved as it was before the transpose).
  const uint64_t* off = g->transposed ? g->in_off : g->off;
  const uint32_t* nbr = g->transposed ? g->in_nbr : g->nbr;
  const uint64_t* ioff = g->transposed ? g->off : g->in_off;
  const uint32_t* inbr = g->transposed ? g->nbr : g->in_nbr;
  if (format == RG_FMT_SG) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return set_error(RG_EIO, "cannot write " + p);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    const char directed = g->symmetric ? 0 : 1;
    const int64_t m = g->m, n = g->n;
    if (std::fwrite(&directed, 1, 1, f) != 1 || std::fwrite(&m, 8, 1, f) != 1 || std::fwrite(&n, 8, 1, f) != 1)
      return set_error(RG_EIO, "short write: " + p);
    RG_TRY(st.to_file(f, off, sizeof(uint64_t) * (n + 1)));
    RG_TRY(st.to_file(f, nbr, sizeof(uint32_t) * m));
    if (!g->symmetric) {
      RG_TRY(st.to_file(f, ioff, sizeof(uint64_t) * (n + 1)));
      RG_TRY(st.to_file(f, inbr, sizeof(uint32_t) * m));
    }
  } else {
    if (g->m > (int64_t)UINT32_MAX) return set_error(RG_EUNSUPPORTED, "binary .idx offsets are 32-bit");
    FILE* c = std::fopen((p + ".config").c_str(), "wb");
    if (!c) return set_error(RG_EIO, "cannot write " + p + ".config");
    std::fprintf(c, "%lld", (long long)g->n);
    std::fclose(c);
    FILE* a = std::fopen((p + ".adj").c_str(), "wb");
    if (!a) return set_error(RG_EIO, "cannot write " + p + ".adj");
    std::unique_ptr<FILE, int (*)(FILE*)> ga(a, std::fclose);
    RG_TRY(st.to_file(a, nbr, sizeof(uint32_t) * g->m));
    std::vector<uint64_t> o64(g->n + 1);
    RG_TRY(rg::ctx_copy(ctx, o64.data(), off, sizeof(uint64_t) * (g->n + 1), hipMemcpyDeviceToHost));
    std::vector<uint32_t> o32(g->n);
    for (int64_t i = 0; i < g->n; ++i) o32[i] = (uint32_t)o64[i];
    FILE* x = std::fopen((p + ".idx").c_str(), "wb");
    if (!x) return set_error(RG_EIO, "cannot write " + p + ".idx");
    std::unique_ptr<FILE, int (*)(FILE*)> gx(x, std::fclose);
    if (std::fwrite(o32.data(), 4, o32.size(), x) != o32.size()) return set_error(RG_EIO, "short write: " + p);
  }
  const std::string idp = p + ".newids";
  if (new_ids) {
    FILE* f = std::fopen(idp.c_str(), "wb");
    if (!f) return set_error(RG_EIO, "cannot write " + idp);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    if (std::fwrite(new_ids, 4, (size_t)g->n, f) != (size_t)g->n) return set_error(RG_EIO, "short write: " + idp);
  } else {
    std::remove(idp.c_str());   // a stale sidecar would mark this graph degree-sorted
  }
  return RG_OK;
}

}  // extern "C"
