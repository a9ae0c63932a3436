// Internal device-side structures of the radar_graph library (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "rg_common.h"

namespace rg {

constexpr uint32_t kNone = 0xffffffffu;   // UINT_E_MAX of Ligra (parallel.h:121)
constexpr int kWave = 64;                 // CDNA wavefront width

#define RG_HIP(call)                                                              \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      return ::rg::set_error(RG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define RG_TRY(call)            \
  do {                          \
    int _rc = (call);           \
    if (_rc != RG_OK) return _rc; \
  } while (0)

// A device buffer that grows on demand and is reused across calls, so the
// per-round loops allocate nothing after warm-up.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t want);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace rg

namespace rg {
// An instantiated HIP graph of a captured launch sequence and the arguments
// it was captured with (pointers, sizes): replayed while they are equal,
// captured again when they change.
struct GraphCache {
  std::vector<uint64_t> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void release();
};
}  // namespace rg

struct rg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  int num_xcds = 8;
  size_t lds_per_cu = 160 * 1024;
  size_t max_lds_per_block = 64 * 1024;
  // Pinned host word for small device->host readbacks (frontier sizes).
  int64_t* h_counts = nullptr;
  // Scratch pools reused by the engine (grown on demand).
  rg::DevBuf counts;      // int64 counters (queue tails, sums)
  rg::DevBuf scratch[12];  // generic pools for apps / kernels ([10], [11]: vertexFilter and reductions)
  rg::DevBuf fr_ids[2];   // frontier id lists (uint32[n] each)
  rg::DevBuf fr_dense;    // frontier byte map (uint8[n]); all zero between uses
  // Optional HIP-event timing of the edge-map launches (rg_ctx_timing).
  bool timing = false;
  struct EvPair { hipEvent_t e[2]; };
  std::vector<EvPair> ev_pairs;
  size_t ev_used = 0;
  // Dense BFS / BC levels: frontier bitmap + per-unit partials; the first
  // bc_bits_zero words of the bitmap are all zero between levels.
  rg::DevBuf bc_dense;
  int64_t bc_bits_zero = 0;
  // Multi-source BFS state (opaque, msbfs.hip), grown on demand.
  void* ms = nullptr;
  // Batched BC state (opaque, bcbatch.hip), grown on demand.
  void* bcb = nullptr;
  // Iteration loops replayed as HIP graphs (rg_ctx_graphs; on by default,
  // not used while per-launch timing is on).
  bool graphs = true;
  // Single-source pull BFS with its level kinds decided on the device
  // (RG_GRAPHS_BFS_LEVELS; off by default).
  bool bfs_device_levels = false;
  rg::GraphCache pr_graph;   // blocked PageRank's maxiters iterations
  rg::GraphCache ms_graph[3];   // a chunk of bit-parallel batch levels, per word width (1, 2, 4)
  rg::GraphCache bfs_graph;     // a chunk of device-controlled single-source BFS levels
  rg::DevBuf bfs_ctl;           // its level control block (BfsCtl)
  // Device array holding the last PageRank ranks / BC dependencies computed
  // on this context (n doubles; valid until the next call on the context).
  const double* last_result = nullptr;
  // Device address of this GPU's fixed scratch array (rg_bfs_noalloc),
  // looked up once per device when the first context is created.
  void* fixed_scratch = nullptr;
};

namespace rg {
// HUBDUP's hubs for one orientation and k: the top-k in-degree vertices as a
// bitmap with a per-32-bit-word rank prefix, rank -> slot and slot -> vertex
// tables (slots by descending in-degree) and the LDS membership filter.
struct HubMeta {
  uint32_t* bits = nullptr;
  uint32_t* rank = nullptr;
  uint32_t* ids = nullptr;       // slot -> vertex
  uint32_t* slot_of = nullptr;   // id-order hub rank -> slot
  uint32_t* filter = nullptr;    // kHubFilterWords words: a bit per hub_hash of each hub
  int64_t k = 0;                 // hubs selected
  int64_t k_req = -1;            // k asked for (-1: empty entry)
  const uint64_t* for_in = nullptr;   // in-offsets of the orientation
  uint64_t used = 0;             // last use (the graph's hub_clock)
};
}  // namespace rg

struct rg_graph {
  rg_ctx* ctx = nullptr;
  int64_t n = 0, m = 0;
  bool symmetric = false;
  // Current orientation: out-CSR and in-CSR (the two swap on transpose).
  uint64_t* off = nullptr;
  uint32_t* nbr = nullptr;
  uint64_t* in_off = nullptr;
  uint32_t* in_nbr = nullptr;
  bool transposed = false;
  bool degree_sorted = false;   // ids by descending degree (made by rg_graph_degree_sort)
  // Hub metadata (HUBDUP), the two most recently used (orientation, k)
  // pairs: BC's forward and backward passes, or the two orientations of a
  // directed graph, each keep theirs.
  rg::HubMeta hubs[2];
  uint64_t hub_clock = 0;
  // Static heavy-chunk list for all-vertices frontiers (PageRank), built
  // for one (out-CSR, heavy degree) pair.
  uint64_t* heavy_items = nullptr;
  int64_t heavy_items_count = 0;
  int heavy_items_deg = -1;
  const uint64_t* heavy_items_off = nullptr;
  // Blocked in-edge layouts (opaque BlockedCSC): PageRank's (2^14-vertex
  // blocks, with partial sums) and the multi-source BFS's (2^12).
  void* blocked = nullptr;
  void* blocked_ms = nullptr;     // of the in-edges of the original orientation
  void* blocked_ms_t = nullptr;   // of the transposed orientation (directed graphs)
  // 1 + the last vertex with in-edges in the orientation whose in-offsets
  // are nz_for (the bound of bottom-up scans).
  const uint64_t* nz_for = nullptr;
  int64_t nz_end = 0;
  // First in-neighbour of every vertex (kNone without in-edges) in the
  // orientation whose in-neighbour array is first_for: the first probe of a
  // bottom-up scan, read coalesced instead of one list line per vertex.
  uint32_t* first_in = nullptr;
  const uint32_t* first_for = nullptr;
  // A view shares the device arrays and cached layouts of `base` (bound to
  // another context, e.g. one serving slot) and frees only what it builds.
  bool view = false;
  const rg_graph* base = nullptr;
};

namespace rg {

// True when p, an array or layout a view holds, is its base's: the view
// must neither free it nor rebuild it in place (other views may be reading
// it), and builds its own instead.
inline bool borrowed(const rg_graph* g, const void* p) {
  if (!g->view || !g->base || !p) return false;
  const rg_graph* b = g->base;
  for (const HubMeta& h : b->hubs)
    if (p == h.bits || p == h.rank || p == h.ids || p == h.slot_of || p == h.filter) return true;
  return p == b->heavy_items || p == b->first_in || p == b->blocked || p == b->blocked_ms ||
         p == b->blocked_ms_t;
}

// Frees an entry's arrays (those the graph does not borrow) and empties it.
void hub_meta_release(rg_graph* g, HubMeta& h);

// Device frontier (Ligra vertexSubset, vertexSubset.h:113-229): a sparse id
// list (always kept) and, for dense rounds, a byte map over all n.  A view:
// the storage belongs to the context pools (ctx_frontiers).
struct Frontier {
  uint32_t* ids = nullptr;  // sparse list
  uint8_t* dense = nullptr; // byte map, all zero outside a dense round
  int64_t size = 0;         // host copy of the number of vertices
  bool all = false;         // the subset is every vertex (ids not materialised)
};

// Points a and b at the context's two frontier lists and its byte map,
// growing them to n entries (the byte map is zeroed when it grows).
int ctx_frontiers(rg_ctx* ctx, int64_t n, Frontier& a, Frontier& b);

int ctx_read_count(rg_ctx* ctx, const int64_t* dev, int64_t* host, int nwords = 1);

// Host<->device copy ordered on the context stream (the stream is
// non-blocking, so a plain hipMemcpy on the null stream would not wait for
// work queued on it); returns after the copy has completed.
int ctx_copy(rg_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);

// With timing on (rg_ctx_timing), records the start event of one edge-map
// launch on the context stream and returns the event pair; the caller
// records element [1] after the launch.  nullptr when timing is off.
hipEvent_t* timing_begin(rg_ctx* ctx);

// Sorts every adjacency list of the CSR (off, nbr) into `sorted` (m
// entries, ascending ids), on the context stream (reorder.hip); chunk_arcs
// > 0 forces the chunked form that graphs of 2^31 arcs or more take.
int sort_lists(rg_ctx* ctx, const uint64_t* off, const uint32_t* nbr, int64_t n, int64_t m, uint64_t chunk_arcs,
               DevBuf& tmp, uint32_t* sorted);

// Cached blocked layout of a graph's in-edges (blocked.hip).
void blocked_free(void* p);

// Multi-source BFS state of a context (msbfs.hip) and batched BC state
// (bcbatch.hip): free, and the device bytes they hold.
void ms_state_free(void* p);
void bcb_state_free(void* p);
int64_t ms_state_bytes(const void* p);
int64_t bcb_state_bytes(const void* p);

// Device bytes a context holds in its pools and query states.
int64_t ctx_device_bytes(const rg_ctx* ctx);
// Frees the context's bit-parallel batch states -- which & 1: the batched
// BFS state (and the captured graphs that use it), which & 2: the batched
// BC state; they are rebuilt on the next batch.
void ctx_release_batch_state(rg_ctx* ctx, int which);

}  // namespace rg
