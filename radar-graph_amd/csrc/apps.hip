// The reference's app functors as device operators, the host-side edgeMap
// dispatcher (Ligra's sparse/dense switch) and the app loops (each app's
// Compute body) of the radar_graph library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "blocked.h"
#include "device.h"
#include "engine.h"
#include "engine_host.h"
#include "queries.h"

namespace rg {

// ---------------------------------------------------------------------------
// Operators.  Each mirrors one reference functor; citations are to
// push-new-orig/Ligra/apps/ unless noted.

// PR_F (PageRank.C:30-43): p_next[d] += p_curr[s]/outdeg(s).  contrib[s]
// holds p_curr[s]/outdeg(s), computed once per vertex by the vertex phase
// (the same IEEE quotient the reference recomputes per arc).
struct PrOp {
  using SrcT = double;
  using HubT = double;
  static constexpr bool kOutput = false;
  static constexpr bool kL2 = true;
  const double* contrib;
  double* acc;
  const int64_t* done;
  __device__ bool skip() const { return *done != 0; }
  __device__ double load_src(uint32_t s) const { return contrib[s]; }
  __device__ bool cond(uint32_t) const { return true; }
  __device__ bool pre_hub(double, uint32_t) const { return true; }
  __device__ bool update(uint32_t, double sv, uint32_t d) const {
    unsafeAtomicAdd(acc + d, sv);
    return false;
  }
  __device__ double hub_identity() const { return 0.0; }
  __device__ void hub_update(double* slot, uint32_t, double sv, uint32_t) const { atomicAdd(slot, sv); }
  __device__ bool hub_flush(uint32_t d, double v) const {
    unsafeAtomicAdd(acc + d, v);
    return false;
  }
  // The per-XCD copy is meant for one XCD's workgroups, but many
  // workgroups add into it: agent scope, the scope that makes the add atomic
  // across them (on this part every global FP64 atomic executes past the L2
  // whatever its scope, so the tier saves no memory traffic either way).
  __device__ void l2_update(double* slot, double sv) const {
    __hip_atomic_fetch_add(slot, sv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ double hub_combine(double x, double y) const { return x + y; }
};

// PageRankDelta's sparse rounds: PR_F's add, and d joins the round's list
// of touched vertices the first time it is reached (a per-vertex stamp of
// the round), so the vertexFilter can run over those alone.  Hub copies in
// LDS only: the per-XCD merge reports no joins.
struct PrdTouchOp {
  using SrcT = double;
  using HubT = double;
  static constexpr bool kOutput = true;
  static constexpr bool kL2 = false;
  const double* contrib;
  double* acc;
  uint32_t* stamp;
  uint32_t tag;
  __device__ bool skip() const { return false; }
  __device__ double load_src(uint32_t s) const { return contrib[s]; }
  __device__ bool cond(uint32_t) const { return true; }
  __device__ bool pre_hub(double, uint32_t) const { return true; }
  // A plain read first: once d carries this round's tag (most arcs into a
  // hub) no returning atomic is issued; a stale read only costs the exchange,
  // which still picks exactly one first toucher.
  __device__ bool first(uint32_t d) const { return stamp[d] != tag && atomicExch(stamp + d, tag) != tag; }
  __device__ bool update(uint32_t, double sv, uint32_t d) const {
    unsafeAtomicAdd(acc + d, sv);
    return first(d);
  }
  __device__ double hub_identity() const { return 0.0; }
  __device__ void hub_update(double* slot, uint32_t, double sv, uint32_t) const { atomicAdd(slot, sv); }
  __device__ bool hub_flush(uint32_t d, double v) const {
    unsafeAtomicAdd(acc + d, v);
    return first(d);
  }
  __device__ void l2_update(double*, double) const {}
  __device__ double hub_combine(double x, double y) const { return x + y; }
};

// BFS_F (BFS.C:28-49): first CAS on Parents[d] wins and d joins the next
// frontier.  Hub copies keep one candidate parent per hub per workgroup
// (hubdup BFS.C:46-68, 72-97).
struct BfsOp {
  using SrcT = uint32_t;
  using HubT = uint32_t;
  static constexpr bool kOutput = true;
  static constexpr bool kL2 = false;
  __device__ void l2_update(uint32_t*, uint32_t) const {}
  __device__ uint32_t hub_combine(uint32_t x, uint32_t) const { return x; }
  uint32_t* parents;
  __device__ bool skip() const { return false; }
  __device__ uint32_t load_src(uint32_t s) const { return s; }
  __device__ bool cond(uint32_t) const { return true; }
  __device__ bool pre_hub(uint32_t, uint32_t d) const { return parents[d] == kNone; }
  __device__ bool update(uint32_t s, uint32_t, uint32_t d) const {
    if (parents[d] != kNone) return false;
    return atomicCAS(parents + d, kNone, s) == kNone;
  }
  __device__ uint32_t hub_identity() const { return kNone; }
  __device__ void hub_update(uint32_t* slot, uint32_t s, uint32_t, uint32_t) const {
    if (*slot == kNone) atomicCAS(slot, kNone, s);
  }
  __device__ bool hub_flush(uint32_t d, uint32_t cand) const {
    if (parents[d] != kNone) return false;
    return atomicCAS(parents + d, kNone, cand) == kNone;
  }
};

// Radii_F (Radii.C:35-64): OR the source's visited word into NextVisited[d]
// when it adds bits; the first such d per round records radii[d] = round.
struct RadiiOp {
  using SrcT = unsigned long long;
  using HubT = unsigned long long;
  static constexpr bool kOutput = true;
  static constexpr bool kL2 = false;
  __device__ void l2_update(unsigned long long*, unsigned long long) const {}
  __device__ unsigned long long hub_combine(unsigned long long x, unsigned long long y) const { return x | y; }
  const unsigned long long* vis;
  unsigned long long* nvis;
  int32_t* radii;
  int32_t round;
  __device__ bool skip() const { return false; }
  __device__ unsigned long long load_src(uint32_t s) const { return vis[s]; }
  __device__ bool cond(uint32_t) const { return true; }
  __device__ bool pre_hub(unsigned long long sv, uint32_t d) const { return (vis[d] | sv) != vis[d]; }
  __device__ bool claim(uint32_t d) const {
    int32_t old = radii[d];
    if (old == round) return false;
    return atomicCAS(radii + d, old, round) == old;
  }
  __device__ bool update(uint32_t, unsigned long long sv, uint32_t d) const {
    unsigned long long vd = vis[d];
    unsigned long long tw = vd | sv;
    if (vd == tw) return false;
    atomicOr(nvis + d, tw);
    return claim(d);
  }
  __device__ unsigned long long hub_identity() const { return 0ull; }
  __device__ void hub_update(unsigned long long* slot, uint32_t, unsigned long long sv,
                             uint32_t d) const {
    unsigned long long vd = vis[d];
    unsigned long long tw = vd | sv;
    if (vd != tw) atomicOr(slot, tw);
  }
  __device__ bool hub_flush(uint32_t d, unsigned long long v) const {
    atomicOr(nvis + d, v);
    return claim(d);
  }
};

// BC_F (BC-manyiter-v2.C:32-56): NumPaths[d] += NumPaths[s] for unvisited
// d; d joins when the add saw 0.
struct BcFwdOp {
  using SrcT = double;
  using HubT = double;
  static constexpr bool kOutput = true;
  static constexpr bool kL2 = false;
  __device__ void l2_update(double*, double) const {}
  __device__ double hub_combine(double x, double y) const { return x + y; }
  double* np;
  const uint8_t* visited;
  __device__ bool skip() const { return false; }
  __device__ double load_src(uint32_t s) const { return np[s]; }
  __device__ bool cond(uint32_t d) const { return visited[d] == 0; }
  __device__ bool pre_hub(double, uint32_t) const { return true; }
  __device__ bool update(uint32_t, double sv, uint32_t d) const {
    return unsafeAtomicAdd(np + d, sv) == 0.0;
  }
  __device__ double hub_identity() const { return 0.0; }
  __device__ void hub_update(double* slot, uint32_t, double sv, uint32_t) const { atomicAdd(slot, sv); }
  __device__ bool hub_flush(uint32_t d, double v) const { return unsafeAtomicAdd(np + d, v) == 0.0; }
};

// BC_Back_F (BC-manyiter-v2.C:58-79): Dependencies[d] += Dependencies[s].
struct BcBackOp {
  using SrcT = double;
  using HubT = double;
  static constexpr bool kOutput = false;
  static constexpr bool kL2 = true;
  // The per-XCD copy is meant for one XCD's workgroups, but many
  // workgroups add into it: agent scope, the scope that makes the add atomic
  // across them (on this part every global FP64 atomic executes past the L2
  // whatever its scope, so the tier saves no memory traffic either way).
  __device__ void l2_update(double* slot, double sv) const {
    __hip_atomic_fetch_add(slot, sv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ double hub_combine(double x, double y) const { return x + y; }
  double* deps;
  const uint8_t* visited;
  __device__ bool skip() const { return false; }
  __device__ double load_src(uint32_t s) const { return deps[s]; }
  __device__ bool cond(uint32_t d) const { return visited[d] == 0; }
  __device__ bool pre_hub(double, uint32_t) const { return true; }
  __device__ bool update(uint32_t, double sv, uint32_t d) const {
    unsafeAtomicAdd(deps + d, sv);
    return false;
  }
  __device__ double hub_identity() const { return 0.0; }
  __device__ void hub_update(double* slot, uint32_t, double sv, uint32_t) const { atomicAdd(slot, sv); }
  __device__ bool hub_flush(uint32_t d, double v) const {
    unsafeAtomicAdd(deps + d, v);
    return false;
  }
};

// ---------------------------------------------------------------------------
// Small vertex kernels.

__global__ void k_scatter_u8(const uint32_t* ids, int64_t count, uint8_t* dst, uint8_t val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[ids[i]] = val;
}

__global__ void k_degsum(const uint32_t* ids, int64_t count, const uint64_t* off, int64_t* out) {
  uint64_t s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = ids[i];
    s += off[v + 1] - off[v];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (__lane_id() == 0 && s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

// Static heavy list for the all-vertices frontier: every vertex with degree
// > heavy_deg contributes ceil(deg / kHeavyChunk) items.
__global__ void k_heavy_list(const uint64_t* off, int64_t n, int heavy_deg, uint64_t* items,
                             int64_t* count) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t dd = off[v + 1] - off[v];
    if (dd > (uint64_t)heavy_deg) {
      int64_t nch = (int64_t)((dd + kHeavyChunk - 1) / kHeavyChunk);
      int64_t base = atomicAdd((unsigned long long*)count, (unsigned long long)nch);
      for (int64_t k = 0; k < nch; ++k) items[base + k] = ((uint64_t)v << 32) | (uint64_t)k;
    }
  }
}

// Builds (once per out-CSR and heavy degree) the chunk list of the vertices
// whose degree exceeds heavy_deg, for all-vertices frontiers.
int static_heavy_list(rg_graph* g, int heavy_deg) {
  if (g->heavy_items && g->heavy_items_deg == heavy_deg && g->heavy_items_off == g->off) return RG_OK;
  rg_ctx* ctx = g->ctx;
  if (g->heavy_items && !borrowed(g, g->heavy_items)) (void)hipFree(g->heavy_items);
  g->heavy_items = nullptr;
  int64_t cap = g->m / kHeavyChunk + g->m / std::max(1, heavy_deg) + 64;
  RG_HIP(hipMalloc(&g->heavy_items, sizeof(uint64_t) * cap));
  int64_t* cnt = ctx->counts.as<int64_t>() + 24;
  RG_HIP(hipMemsetAsync(cnt, 0, sizeof(int64_t), ctx->stream));
  hipLaunchKernelGGL(k_heavy_list, grid_for(ctx, g->n), 256, 0, ctx->stream, g->off, g->n, heavy_deg,
                     g->heavy_items, cnt);
  RG_HIP(hipGetLastError());
  RG_TRY(ctx_read_count(ctx, cnt, &g->heavy_items_count, 1));
  g->heavy_items_deg = heavy_deg;
  g->heavy_items_off = g->off;
  return RG_OK;
}

// Ligra edgeMapData (ligra.h:264-302) on the device: chooses the dense
// (edgeMapDenseForward) or sparse (edgeMapSparse) traversal from |F| plus
// the frontier's out-degree sum, runs the push, and (for operators with
// output) appends the next frontier to out_ids with its size and degree sum
// in out_info[0..1] (which the caller zeroes).  A frontier with `all` set is
// every vertex (PageRank's frontier, PageRank.C:112-115).
template <class Op>
static int edge_map(const Engine& E, const Op& op, Frontier& F, int64_t f_degsum,
                    uint32_t* out_ids, int64_t* out_info) {
  rg_graph* g = E.g;
  if (!F.all && F.size == 0) return RG_OK;
  RG_HIP(hipMemsetAsync(E.heavy_count, 0, sizeof(int64_t), E.ctx->stream));
  EdgeArgs a{};
  a.off = g->off;
  a.nbr = g->nbr;
  a.out_ids = out_ids;
  a.out_info = out_info;
  a.heavy = E.heavy;
  a.heavy_count = E.heavy_count;
  a.heavy_deg = E.heavy_deg;
  a.hub_k = E.hub_k;
  a.hub_k2 = Op::kL2 ? E.hub_k2 : E.hub_k;
  if (E.hm) {
    a.hub_bits = E.hm->bits;
    a.hub_rank = E.hm->rank;
    a.hub_slot_of = E.hm->slot_of;
    a.hub_ids = E.hm->ids;
    a.hub_filter = E.hm->filter;
  }
  a.l2_copies = E.l2_copies;
  a.l2_ncopies = E.l2_ncopies;
  if (F.all) {
    a.count = g->n;
    RG_TRY(static_heavy_list(g, E.heavy_deg));
    a.static_heavy = g->heavy_items;
    a.static_heavy_count = g->heavy_items_count;
    return launch_hub<Op, SRC_ALL>(E, op, a);
  }
  if (f_degsum == 0) return RG_OK;   // ligra.h:287
  if (E.hub_mode != HUB_NONE && F.size + f_degsum < E.hub_min_arcs) {
    Engine plain = E;
    plain.hub_mode = HUB_NONE;
    return edge_map(plain, op, F, f_degsum, out_ids, out_info);
  }
  if (F.size + f_degsum > E.threshold) {
    const int grid = grid_for(E.ctx, F.size);
    hipLaunchKernelGGL(k_scatter_u8, grid, 256, 0, E.ctx->stream, F.ids, F.size, F.dense, (uint8_t)1);
    a.count = g->n;
    a.dense = F.dense;
    RG_TRY((launch_hub<Op, SRC_DENSE>(E, op, a)));
    hipLaunchKernelGGL(k_scatter_u8, grid, 256, 0, E.ctx->stream, F.ids, F.size, F.dense, (uint8_t)0);
    RG_HIP(hipGetLastError());
    return RG_OK;
  }
  a.count = F.size;
  a.ids = F.ids;
  // The light kernel spreads each wave's 64 vertices' arcs over its lanes, so
  // a small frontier with long lists (BFS level 2 on RMAT-26: ~700 vertices,
  // ~20M arcs, 12 waves for 1.2 ms) leaves the device idle: below 16 light
  // waves per CU, lists longer than a wave become chunk items of the heavy
  // kernel, one wave each, when the queue holds them (at most |F| +
  // degsum / kHeavyChunk items).
  if (F.size < (int64_t)E.ctx->num_cus * 16 * kWave && E.heavy_deg > kWave &&
      F.size + f_degsum / kHeavyChunk + 64 <= E.heavy_cap)
    a.heavy_deg = kWave;
  return launch_hub<Op, SRC_LIST>(E, op, a);
}

// Sum of out-degrees of a frontier list (device), read back.
static int frontier_degsum(rg_ctx* ctx, const rg_graph* g, const uint32_t* ids, int64_t count,
                           int64_t* out, int64_t* counter = nullptr) {
  int64_t* d = counter ? counter : ctx->counts.as<int64_t>() + 8;
  RG_HIP(hipMemsetAsync(d, 0, sizeof(int64_t), ctx->stream));
  if (count > 0)
    hipLaunchKernelGGL(k_degsum, grid_for(ctx, count), 256, 0, ctx->stream, ids, count, g->off, d);
  RG_HIP(hipGetLastError());
  return ctx_read_count(ctx, d, out, 1);
}

// ---------------------------------------------------------------------------
// Hub metadata.

__global__ void k_in_degree(const uint64_t* in_off, int64_t n, uint32_t* deg) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    deg[i] = (uint32_t)(in_off[i + 1] - in_off[i]);
}

// HUBDUP's identifyHubs (hubdup Ligra/ligra/IO.h:726-804): hubs are the
// vertices whose in-degree exceeds the k-th largest in-degree (0-based),
// so at most k of them.  Stored as a bitmap with a per-32-bit-word rank
// prefix (id-order rank of a hub), a rank -> slot table and a slot -> vertex
// table; slots are ordered by descending in-degree so the first slots (the
// LDS tier) are the heaviest hubs.
static int build_hub_bitmap(rg_graph* g, int64_t k, const HubMeta** out) {
  rg_ctx* ctx = g->ctx;
  for (HubMeta& h : g->hubs)
    if (h.k_req == k && h.for_in == g->in_off && h.bits) {
      h.used = ++g->hub_clock;
      *out = &h;
      return RG_OK;
    }
  HubMeta& hm = g->hubs[0].used <= g->hubs[1].used ? g->hubs[0] : g->hubs[1];   // least recently used
  hub_meta_release(g, hm);
  const int64_t n = g->n;
  std::vector<uint32_t> deg(n);
  DevBuf tmp;
  RG_TRY(tmp.ensure(sizeof(uint32_t) * n));
  hipLaunchKernelGGL(k_in_degree, grid_for(ctx, n), 256, 0, ctx->stream, g->in_off, n, tmp.as<uint32_t>());
  RG_HIP(hipMemcpyAsync(deg.data(), tmp.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(hipStreamSynchronize(ctx->stream));
  tmp.release();
  uint32_t thresh;
  if (k >= n) {
    thresh = 0;   // every vertex with an in-edge
  } else {
    std::vector<uint32_t> s(deg);
    std::nth_element(s.begin(), s.begin() + k, s.end(), std::greater<uint32_t>());
    thresh = s[k];
  }
  const int64_t nw = (n + 31) / 32;
  std::vector<uint32_t> bits(nw, 0), rank(nw, 0), ids;
  for (int64_t v = 0; v < n; ++v)
    if (deg[v] > thresh) {
      bits[v >> 5] |= 1u << (v & 31);
      ids.push_back((uint32_t)v);
    }
  uint32_t acc = 0;
  for (int64_t w = 0; w < nw; ++w) {
    rank[w] = acc;
    acc += (uint32_t)__builtin_popcount(bits[w]);
  }
  // Slots by (in-degree, id) descending; slot_of maps id-order rank -> slot.
  std::vector<uint32_t> by_deg(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) by_deg[i] = (uint32_t)i;
  std::sort(by_deg.begin(), by_deg.end(), [&](uint32_t x, uint32_t y) {
    uint32_t dx = deg[ids[x]], dy = deg[ids[y]];
    return dx != dy ? dx > dy : ids[x] > ids[y];
  });
  std::vector<uint32_t> slot_of(ids.size()), slot_ids(ids.size());
  for (size_t s = 0; s < by_deg.size(); ++s) {
    slot_of[by_deg[s]] = (uint32_t)s;
    slot_ids[s] = ids[by_deg[s]];
  }
  std::vector<uint32_t> filt(kHubFilterWords, 0);
  for (uint32_t v : ids) {
    const uint32_t h = hub_hash(v);
    filt[h >> 5] |= 1u << (h & 31);
  }
  const size_t nh = std::max<size_t>(1, ids.size());
  RG_HIP(hipMalloc(&hm.bits, sizeof(uint32_t) * nw));
  RG_HIP(hipMalloc(&hm.rank, sizeof(uint32_t) * nw));
  RG_HIP(hipMalloc(&hm.ids, sizeof(uint32_t) * nh));
  RG_HIP(hipMalloc(&hm.slot_of, sizeof(uint32_t) * nh));
  RG_HIP(hipMalloc(&hm.filter, sizeof(uint32_t) * kHubFilterWords));
  RG_TRY(rg::ctx_copy(ctx, hm.filter, filt.data(), sizeof(uint32_t) * kHubFilterWords, hipMemcpyHostToDevice));
  RG_TRY(rg::ctx_copy(ctx, hm.bits, bits.data(), sizeof(uint32_t) * nw, hipMemcpyHostToDevice));
  RG_TRY(rg::ctx_copy(ctx, hm.rank, rank.data(), sizeof(uint32_t) * nw, hipMemcpyHostToDevice));
  if (!ids.empty()) {
    RG_TRY(rg::ctx_copy(ctx, hm.ids, slot_ids.data(), sizeof(uint32_t) * ids.size(), hipMemcpyHostToDevice));
    RG_TRY(rg::ctx_copy(ctx, hm.slot_of, slot_of.data(), sizeof(uint32_t) * ids.size(), hipMemcpyHostToDevice));
  }
  hm.k = (int64_t)ids.size();
  hm.k_req = k;
  hm.for_in = g->in_off;
  hm.used = ++g->hub_clock;
  *out = &hm;
  return RG_OK;
}

// Resolves variant/opts into an Engine.  elem = bytes of one hub copy entry.
int make_engine(rg_graph* g, const rg_opts* o, size_t elem, Engine& E, int wgs_per_cu, bool l2_tier,
                uint64_t* heavy, int64_t heavy_cap, int64_t* heavy_count) {
  rg_ctx* ctx = g->ctx;
  E.ctx = ctx;
  E.g = g;
  E.threshold = (o && o->threshold >= 0) ? o->threshold : (int64_t)(int32_t)(g->m / 20);
  E.block = (o && o->wg_size >= 1024) ? 1024 : 256;
  E.heavy_deg = (o && o->heavy_degree > 0) ? o->heavy_degree : 1024;
  E.hub_mode = HUB_NONE;
  E.hub_k = E.hub_k2 = 0;
  E.l2_copies = nullptr;
  E.l2_ncopies = 0;
  const int variant = o ? o->variant : RG_PUSH;
  if (variant == RG_HUBDUP || variant == RG_RADAR) {
    const int wpb = E.block / kWave;
    size_t static_lds = wpb * ((kWave + 1) * 4 + kWave * (8 + 4 + 8) + kQueue * 4);
    if (variant == RG_HUBDUP) static_lds += sizeof(uint32_t) * kHubFilterWords;   // the membership filter
    size_t budget = ctx->max_lds_per_block > static_lds ? ctx->max_lds_per_block - static_lds : 0;
    int64_t kmax = (int64_t)(budget / elem);
    // LDS tier: the requested k, or by default what leaves room for
    // wgs_per_cu workgroups per CU (the reference sizes its copies from the
    // cache budget llcCap, hubdup IO.h:749-755).
    const size_t share = ctx->lds_per_cu / (size_t)std::max(1, wgs_per_cu);
    const int64_t kdef = share > static_lds ? std::min<int64_t>(kmax, (int64_t)((share - static_lds) / elem)) : 0;
    int64_t k = (o && o->hubs > 0) ? std::min<int64_t>(o->hubs, kmax) : kdef;
    k = std::min<int64_t>(k, g->n);
    // Per-XCD tier, when asked for: its adds execute past the L2 like any
    // global atomic on this part, and the hub sweep measured no gain from it.
    int64_t k2 = k;
    if (l2_tier && elem == sizeof(double) && o && o->l2_hubs > k) k2 = std::min<int64_t>(o->l2_hubs, g->n);
    if (k > 0) {
      if (variant == RG_RADAR) {
        E.hub_mode = HUB_PREFIX;
        E.hub_k = k;
        E.hub_k2 = k2;
      } else {
        RG_TRY(build_hub_bitmap(g, k2, &E.hm));
        E.hub_mode = HUB_BITMAP;
        E.hub_k2 = E.hm->k;
        E.hub_k = std::min<int64_t>(k, E.hm->k);
        if (E.hub_k2 == 0) E.hub_mode = HUB_NONE;
      }
    }
    if (E.hub_mode != HUB_NONE && E.hub_k2 > E.hub_k) {
      E.l2_ncopies = ctx->num_xcds;
      size_t bytes = (size_t)E.l2_ncopies * (size_t)E.hub_k2 * elem;
      DevBuf& b = ctx->scratch[6];
      RG_TRY(b.ensure(bytes));
      RG_HIP(hipMemsetAsync(b.p, 0, bytes, ctx->stream));
      E.l2_copies = b.p;
    }
  }
  // Small launches skip the hub copies (o->hub_min_arcs: < 0 never, 0 the
  // default of 4096 frontier arcs per CU).
  E.hub_min_arcs = (o && o->hub_min_arcs < 0) ? 0
                   : (o && o->hub_min_arcs > 0) ? o->hub_min_arcs : (int64_t)ctx->num_cus * 4096;
  // Heavy queue: at most m/heavy_deg heavy vertices, each split in chunks.
  int64_t cap = heavy_queue_items(g->m, E.heavy_deg);
  if (heavy) {
    if (heavy_cap < cap || !heavy_count) return set_error(RG_EINVAL, "edgeMap: heavy queue too small");
    E.heavy = heavy;
    E.heavy_cap = heavy_cap;
    E.heavy_count = heavy_count;
    return RG_OK;
  }
  RG_TRY(ctx->scratch[7].ensure(sizeof(uint64_t) * cap));
  E.heavy = ctx->scratch[7].as<uint64_t>();
  E.heavy_cap = cap;
  E.heavy_count = ctx->counts.as<int64_t>() + 0;
  return RG_OK;
}

// ---------------------------------------------------------------------------
// PageRank (PageRank.C:92-141).

__global__ void k_pr_init(const uint64_t* off, int64_t n, double* rank, double* contrib, double* acc,
                          double one_over_n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t deg = off[v + 1] - off[v];
    rank[v] = one_over_n;
    contrib[v] = deg ? one_over_n / (double)(uint32_t)deg : 0.0;
    acc[v] = 0.0;
  }
}

// Fused vertex phase: PR_Vertex_F (p_next = 0.85 p_next + 0.15/n), the L1
// term |p_curr - p_next|, PR_Vertex_Reset and the swap, plus the next
// iteration's per-source quotient.  The multiply and add are rounded
// separately, as the reference's host code does.
__global__ __launch_bounds__(256) void k_pr_vertex(const uint64_t* off, int64_t n, double* rank,
                                                   double* contrib, double* acc, double damping,
                                                   double added, double* partial,
                                                   const int64_t* done) {
  if (*done) return;
  double diff = 0.0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    double nw = __dadd_rn(__dmul_rn(damping, acc[v]), added);
    diff += fabs(rank[v] - nw);
    rank[v] = nw;
    uint64_t deg = off[v + 1] - off[v];
    contrib[v] = deg ? nw / (double)(uint32_t)deg : 0.0;
    acc[v] = 0.0;
  }
  __shared__ double red[256];
  red[threadIdx.x] = diff;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Sums the block partials in a fixed order; records L1 for this iteration
// and, the first time it drops below epsilon, the (1-based) iteration.
__global__ void k_pr_finish(const double* partial, int nparts, double* l1_hist, int64_t iter,
                            int64_t* done, double epsilon) {
  if (*done) return;
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    l1_hist[iter - 1] = red[0];
    if (red[0] < epsilon) *done = iter;
  }
}

int pagerank_push(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                  int64_t* iters_out, double* l1_out, double* ms_out);
int pagerank_pull(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                  int64_t* iters_out, double* l1_out, double* ms_out);
int pagerank_blocked(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                     int64_t* iters_out, double* l1_out, double* ms_out);

int pagerank_push(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                  int64_t* iters_out, double* l1_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  Engine E;
  RG_TRY(make_engine(g, o, sizeof(double), E, 1, true));
  DevBuf &b_rank = ctx->scratch[0], &b_contrib = ctx->scratch[1], &b_acc = ctx->scratch[2],
         &b_part = ctx->scratch[3];
  RG_TRY(b_rank.ensure(sizeof(double) * n));
  RG_TRY(b_contrib.ensure(sizeof(double) * n));
  RG_TRY(b_acc.ensure(sizeof(double) * n));
  const int vgrid = std::min<int>(grid_for(ctx, n), 1024);
  RG_TRY(b_part.ensure(sizeof(double) * (vgrid + std::max<int64_t>(maxiters, 1) + 8)));
  double* rank = b_rank.as<double>();
  double* contrib = b_contrib.as<double>();
  double* acc = b_acc.as<double>();
  double* partial = b_part.as<double>();
  double* l1_hist = partial + vgrid;
  int64_t* done = ctx->counts.as<int64_t>() + 16;
  RG_HIP(hipMemsetAsync(done, 0, sizeof(int64_t), ctx->stream));
  const double damping = 0.85, epsilon = 0.0000001;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * (1 / (double)n);
  hipLaunchKernelGGL(k_pr_init, grid_for(ctx, n), 256, 0, ctx->stream, g->off, n, rank, contrib, acc,
                     one_over_n);
  RG_HIP(hipGetLastError());
  Frontier all;
  all.all = true;
  PrOp op{contrib, acc, done};
  EventTimer tm(ctx->stream);
  for (int64_t it = 1; it <= maxiters; ++it) {
    RG_TRY(edge_map(E, op, all, g->m, nullptr, nullptr));
    hipLaunchKernelGGL(k_pr_vertex, vgrid, 256, 0, ctx->stream, g->off, n, rank, contrib, acc, damping,
                       added, partial, done);
    hipLaunchKernelGGL(k_pr_finish, 1, 256, 0, ctx->stream, partial, vgrid, l1_hist, it, done, epsilon);
    RG_HIP(hipGetLastError());
  }
  double ms = tm.stop_ms();
  int64_t dn = 0;
  RG_TRY(ctx_read_count(ctx, done, &dn, 1));
  int64_t last = dn ? dn : maxiters;
  double l1 = 0.0;
  if (last >= 1)
    RG_TRY(rg::ctx_copy(ctx, &l1, l1_hist + (last - 1), sizeof(double), hipMemcpyDeviceToHost));
  if (iters_out) *iters_out = dn ? dn : maxiters + 1;
  if (l1_out) *l1_out = l1;
  if (ms_out) *ms_out = ms;
  ctx->last_result = rank;
  if (ranks_out) RG_TRY(rg::ctx_copy(ctx, ranks_out, rank, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

// ---------------------------------------------------------------------------
// PageRankDelta (PageRankDelta.C:102-154).

// Delta itself is not stored: the edge phase reads only Delta[s]/outdeg(s),
// kept in `contrib` (1/n over the degree before round 1).
__global__ void k_prd_init(const uint64_t* off, int64_t n, double* p, double* ngh, double* contrib,
                           double one_over_n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t deg = off[v + 1] - off[v];
    p[v] = 0.0;
    ngh[v] = 0.0;
    contrib[v] = deg ? one_over_n / (double)(uint32_t)deg : 0.0;
  }
}

// Fused vertexFilter of one round: PR_Vertex_F_FirstRound (round 1,
// PageRankDelta.C:51-64) or PR_Vertex_F (66-78), the L1 term |Delta|, the
// reset of nghSum (80-88) and the next round's quotient Delta[s]/outdeg(s)
// for active s (0 for inactive ones, so an all-vertices pass adds nothing
// for them).  Active vertices are appended to out_ids with their out-degree
// sum.  nghSum comes from `ngh`, which is reset.  With `list`, only the
// vertices list[0, *list_count) are visited: after a sparse round every
// other vertex has nghSum 0, so Delta 0, p unchanged and contrib 0 (kept
// so: the previous round's active vertices have their contrib cleared
// first).  Multiplies and adds are
// rounded separately, as the reference's host code computes them.
__global__ __launch_bounds__(256) void k_prd_vertex(const uint64_t* off, int64_t n, const uint32_t* list,
                                                    const int64_t* list_count, double* p, double* ngh,
                                                    double* contrib, int first,
                                                    double damping, double added, double one_over_n,
                                                    double eps2, double* partial, uint32_t* out_ids,
                                                    int64_t* out_info) {
  __shared__ uint32_t s_q[4][kQueue];
  __shared__ double red[256];
  EdgeArgs ea{};
  ea.off = off;
  ea.out_ids = out_ids;
  ea.out_info = out_info;
  WaveQueue q{s_q[threadIdx.x / kWave], 0, 0};
  double l1 = 0.0;
  const int64_t total = list ? *list_count : n;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    bool act = false;
    int64_t v = i;
    if (i < total) {
      if (list) v = list[i];
      const double ns = ngh[v];
      ngh[v] = 0.0;
      double pv = p[v], dl;
      if (first) {
        dl = __dadd_rn(__dmul_rn(damping, ns), added);
        pv = __dadd_rn(pv, dl);
        dl = __dsub_rn(dl, one_over_n);
        act = fabs(dl) > __dmul_rn(eps2, pv);
      } else {
        dl = __dmul_rn(ns, damping);
        act = fabs(dl) > __dmul_rn(eps2, pv);
        if (act) pv = __dadd_rn(pv, dl);
      }
      p[v] = pv;
      l1 += fabs(dl);
      const uint64_t deg = off[v + 1] - off[v];
      contrib[v] = (act && deg) ? dl / (double)(uint32_t)deg : 0.0;
    }
    wq_push(q, ea, act, (uint32_t)v);
  }
  wq_finish(q, ea);
  red[threadIdx.x] = l1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_zero_at(const uint32_t* ids, int64_t count, double* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    x[ids[i]] = 0.0;
}

__global__ void k_prd_l1(const double* partial, int nparts, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

int pagerank_delta(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out, int64_t* iters_out,
                   double* l1_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (n == 0) return set_error(RG_EINVAL, "rg_pagerank_delta: empty graph");
  Engine E;
  RG_TRY(make_engine(g, o, sizeof(double), E));
  const bool blocked = o && o->variant == RG_BLOCKED;
  // On a graph in degree order (hubs are ids 0..k-1) the blocked variant's
  // sparse rounds push with RADAR's LDS copies of the hubs' sums: most arcs
  // of a sparse round land on hubs.  By default the copies leave room for
  // four workgroups per CU (~4K hubs).  Symmetric RMAT-26, 20 rounds
  // (stops after 15): 30.65 ms without copies, 29.09 with one workgroup's
  // worth per CU (~18K hubs), 28.62 with 4096 hubs.
  Engine Es;
  const bool hub_sparse = blocked && g->degree_sorted;
  if (hub_sparse) {
    rg_opts os = *o;
    os.variant = RG_RADAR;
    RG_TRY(make_engine(g, &os, sizeof(double), Es, 4));
  }
  DevBuf &b_p = ctx->scratch[0], &b_ngh = ctx->scratch[2], &b_part = ctx->scratch[3], &b_contrib = ctx->scratch[4],
         &b_unit = ctx->scratch[5], &b_contrib2 = ctx->scratch[1];
  RG_TRY(b_p.ensure(sizeof(double) * n));
  RG_TRY(b_ngh.ensure(sizeof(double) * n));
  RG_TRY(b_contrib.ensure(sizeof(double) * n));
  if (blocked) RG_TRY(b_contrib2.ensure(sizeof(double) * n));
  const int vgrid = std::min<int>(grid_for(ctx, n), 1024);
  RG_TRY(b_part.ensure(sizeof(double) * (vgrid + 8)));
  double* p = b_p.as<double>();
  double* ngh = b_ngh.as<double>();
  double* contrib = b_contrib.as<double>();
  // Blocked dense rounds read contrib while they write the next quotients,
  // so they write the other buffer and the two swap.
  double* contrib_alt = blocked ? b_contrib2.as<double>() : nullptr;
  double* partial = b_part.as<double>();
  int64_t* zero = ctx->counts.as<int64_t>() + 32;
  int64_t* info = ctx->counts.as<int64_t>() + 34;   // [count, out-degree sum]
  int64_t* tinfo = ctx->counts.as<int64_t>() + 36;  // touched list of a sparse round
  // The round's L1 norm sits after them, so one read returns count, degree sum and L1.
  double* l1_dev = reinterpret_cast<double*>(ctx->counts.as<int64_t>() + 38);
  RG_HIP(hipMemsetAsync(zero, 0, sizeof(int64_t), ctx->stream));
  // Sparse rounds: the vertices their push reaches, and a per-vertex stamp
  // of the last round that reached each (tags are rounds, from 2).
  DevBuf &b_touch = ctx->scratch[8], &b_stamp = ctx->scratch[9];
  RG_TRY(b_touch.ensure(sizeof(uint32_t) * n));
  RG_TRY(b_stamp.ensure(sizeof(uint32_t) * n));
  uint32_t* touched = b_touch.as<uint32_t>();
  uint32_t* stamp = b_stamp.as<uint32_t>();
  RG_TRY(fill<uint32_t>(ctx, stamp, n, 0u));
  Frontier fa, fb;
  RG_TRY(ctx_frontiers(ctx, n, fa, fb));
  const double damping = 0.85, epsilon = 0.0000001, epsilon2 = 0.01;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * one_over_n;
  hipLaunchKernelGGL(k_prd_init, grid_for(ctx, n), 256, 0, ctx->stream, g->off, n, p, ngh, contrib, one_over_n);
  RG_HIP(hipGetLastError());
  Frontier* F = &fa;
  Frontier* N = &fb;
  F->all = true;
  F->size = n;
  int64_t degsum = g->m;
  double L1 = 0.0;
  int64_t round = 0;
  EventTimer tm(ctx->stream);
  while (round++ < maxiters) {
    const uint32_t* list = nullptr;   // vertexFilter over the touched vertices only
    const bool dense_blocked = blocked && (F->all || F->size + degsum > E.threshold);
    RG_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int64_t), ctx->stream));
    if (dense_blocked) {
      const PrdStep st{damping, added, one_over_n, epsilon2, round == 1 ? 1 : 0};
      RG_TRY(blocked_prd_round(g, o, contrib, contrib_alt, p, st, N->ids, info, l1_dev, zero, b_unit));
      std::swap(contrib, contrib_alt);
    } else if (!F->all && F->size + degsum <= E.threshold) {
      // Sparse push (edge_map's list traversal) that also lists the vertices
      // it reaches; then the previous active vertices' contrib is cleared
      // (the vertex step sets it again for those still active).
      RG_HIP(hipMemsetAsync(tinfo, 0, 2 * sizeof(int64_t), ctx->stream));
      PrdTouchOp top{contrib, ngh, stamp, (uint32_t)round};
      RG_TRY(edge_map(hub_sparse ? Es : E, top, *F, degsum, touched, tinfo));
      if (F->size > 0)
        hipLaunchKernelGGL(k_zero_at, grid_for(ctx, F->size), 256, 0, ctx->stream, F->ids, F->size, contrib);
      list = touched;
    } else {
      PrOp op{contrib, ngh, zero};
      RG_TRY(edge_map(E, op, *F, degsum, nullptr, nullptr));
    }
    if (!dense_blocked) {
      hipLaunchKernelGGL(k_prd_vertex, vgrid, 256, 0, ctx->stream, g->off, n, list, tinfo, p, ngh, contrib,
                         round == 1 ? 1 : 0, damping, added, one_over_n, epsilon2, partial, N->ids, info);
      hipLaunchKernelGGL(k_prd_l1, 1, 256, 0, ctx->stream, partial, vgrid, l1_dev);
    }
    RG_HIP(hipGetLastError());
    int64_t h[5];
    RG_TRY(ctx_read_count(ctx, info, h, 5));
    std::memcpy(&L1, &h[4], sizeof(double));
    if (L1 < epsilon) break;
    N->all = false;
    N->size = h[0];
    degsum = h[1];
    std::swap(F, N);
  }
  const double ms = tm.stop_ms();
  ctx->last_result = p;
  if (iters_out) *iters_out = round;
  if (l1_out) *l1_out = L1;
  if (ms_out) *ms_out = ms;
  if (ranks_out) RG_TRY(rg::ctx_copy(ctx, ranks_out, p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

// ---------------------------------------------------------------------------
// BFS (BFS.C:73-117).

// Sum of level_hash(v, level) over a frontier list whose size is on the
// device (the level checksum of a BFS query).
__global__ void k_level_cs(const uint32_t* ids, const int64_t* count, uint32_t level, unsigned long long* out) {
  const int64_t c = *count;
  unsigned long long s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < c; i += (int64_t)gridDim.x * blockDim.x)
    s += level_hash(ids[i], level);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (__lane_id() == 0 && s) atomicAdd(out, s);
}

// Cost of a top-down arc (a random parent word) against a bottom-up step
// (a sequential parent read or list entry) in the pull engine's choice to
// stay bottom-up.  On a degree-sorted RMAT-26: 3.107 ms per source with the
// rule off (or at 4, where it never fires), 3.07 at 16, 32 and 64.
constexpr int64_t kPullStayArcCost = 16;

// Single-source BFS with the level kinds decided on the device (the pull
// engine's loop): the control block below replaces the host's per-level
// read of the frontier size and degree sum, so a chunk of kBfsLevelsPerSync
// levels -- every level's push kernels (skipped unless the level is
// top-down) and pull kernels (skipped unless bottom-up), the checksum and
// k_bfs_level_end -- is launched, or replayed as a captured HIP graph,
// without a host round trip; the host checks for completion once per chunk.
enum { BFS_DONE = 0, BFS_TOPDOWN = 1, BFS_BOTTOMUP = 2 };
constexpr int kBfsLevelsPerSync = 4;   // even: a chunk ends on the list it started from

struct BfsCtl {
  int64_t mode;                 // BFS_DONE / BFS_TOPDOWN / BFS_BOTTOMUP: this level's kind
  int64_t fcount, fdeg;         // the frontier: size, out-degree sum
  int64_t ncount, ndeg;         // the next frontier (the level's out_info)
  int64_t heavy_count;          // top-down heavy queue items
  int64_t pheavy_count;         // bottom-up heavy vertices
  int64_t reached, edges, iter;
  int64_t chain;                // PullSel::chain of this level
  int64_t heavy_deg;            // top-down wave-per-vertex cut-over of this level
  unsigned long long cs;        // summary checksum (levels >= 1)
  // constants of the query
  int64_t threshold, n, m, heavy_cap, small_list, base_heavy_deg;
};

// The host loop's choice for the frontier in c (bfs_push): bottom-up when
// |F| + degsum passes the threshold, or when the previous level was
// bottom-up and staying costs less than pushing degsum random arcs; else
// top-down, with lists longer than a wave sent to the heavy kernel when
// the frontier fills fewer than 16 light waves per CU.
__device__ void bfs_plan(BfsCtl* c) {
  if (c->fcount == 0) {
    c->mode = BFS_DONE;
    return;
  }
  const bool stay = c->chain != 0 && c->fdeg * kPullStayArcCost > c->n + (c->m - c->edges);
  if (c->fdeg > 0 && (c->fcount + c->fdeg > c->threshold || stay)) {
    c->mode = BFS_BOTTOMUP;
  } else {
    c->mode = BFS_TOPDOWN;
    c->heavy_deg = (c->fcount < c->small_list && c->base_heavy_deg > kWave &&
                    c->fcount + c->fdeg / kHeavyChunk + 64 <= c->heavy_cap)
                       ? kWave
                       : c->base_heavy_deg;
  }
}

__global__ void k_bfs_ctl_init(BfsCtl* c, const uint64_t* off, uint32_t src, int64_t threshold, int64_t n,
                               int64_t m, int64_t heavy_cap, int64_t small_list, int64_t base_heavy_deg) {
  c->fcount = 1;
  c->fdeg = (int64_t)(off[src + 1] - off[src]);
  c->ncount = c->ndeg = c->heavy_count = c->pheavy_count = 0;
  c->reached = 1;
  c->edges = c->fdeg;
  c->iter = 0;
  c->chain = 0;
  c->cs = 0;
  c->threshold = threshold;
  c->n = n;
  c->m = m;
  c->heavy_cap = heavy_cap;
  c->small_list = small_list;
  c->base_heavy_deg = base_heavy_deg;
  c->heavy_deg = base_heavy_deg;
  bfs_plan(c);
}

// After a level: the next frontier becomes the frontier, counters reset,
// the next level's kind chosen.
__global__ void k_bfs_level_end(BfsCtl* c) {
  if (c->mode == BFS_DONE) return;
  c->iter += 1;
  c->reached += c->ncount;
  c->edges += c->ndeg;
  c->chain = c->mode == BFS_BOTTOMUP ? (c->chain == 1 ? 2 : 1) : 0;
  c->fcount = c->ncount;
  c->fdeg = c->ndeg;
  c->ncount = c->ndeg = c->heavy_count = c->pheavy_count = 0;
  bfs_plan(c);
}

__global__ void k_level_cs_ctl(const uint32_t* ids, const BfsCtl* c, unsigned long long* out) {
  if (c->mode == BFS_DONE) return;
  const int64_t cnt = c->ncount;
  const uint32_t level = (uint32_t)(c->iter + 1);
  unsigned long long s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    s += level_hash(ids[i], level);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (__lane_id() == 0 && s) atomicAdd(out, s);
}

// BfsOp whose launches run only on top-down levels.
struct BfsLevelOp : BfsOp {
  const int64_t* mode;
  __device__ bool skip() const { return *mode != BFS_TOPDOWN; }
};

// The level loop of the pull engine run by the device (frontier list
// ids[0] = {src} and parents initialised by the caller).
int bfs_pull_device_loop(rg_graph* g, const Engine& E, const rg_opts* o, uint32_t src, uint32_t* parents,
                         uint32_t* const list[2], bool checksum, int64_t* iters, int64_t* reached, int64_t* edges,
                         uint64_t* cs_out) {
  rg_ctx* ctx = g->ctx;
  hipStream_t st = ctx->stream;
  PullBufs pb;
  RG_TRY(bfs_pull_buffers(g, o ? o->heavy_degree : 0, &pb));
  RG_TRY(ctx->bfs_ctl.ensure(sizeof(BfsCtl)));
  BfsCtl* c = ctx->bfs_ctl.as<BfsCtl>();
  RG_TRY(bfs_pull_zero_tails(ctx, pb));
  hipLaunchKernelGGL(k_bfs_ctl_init, 1, 1, 0, st, c, g->off, src, E.threshold, g->n, g->m, E.heavy_cap,
                     (int64_t)ctx->num_cus * 16 * kWave, (int64_t)E.heavy_deg);
  RG_HIP(hipGetLastError());
  PullSel sel;
  sel.mode = &c->mode;
  sel.chain = &c->chain;
  sel.want = BFS_BOTTOMUP;
  sel.zero_bits = pb.bits;
  sel.full0 = pb.full[0];
  sel.full1 = pb.full[1];
  BfsLevelOp op;
  op.parents = parents;
  op.mode = &c->mode;
  EdgeArgs ea{};
  ea.off = g->off;
  ea.nbr = g->nbr;
  // The list size is read on the device; the bound only sizes the grid (a
  // persistent one of four 256-thread workgroups per CU: top-down levels of
  // this engine are the small frontiers).
  ea.count = (int64_t)ctx->num_cus * 4 * 256;
  ea.count_dev = &c->fcount;
  ea.out_info = &c->ncount;
  ea.heavy = E.heavy;
  ea.heavy_count = &c->heavy_count;
  ea.heavy_deg = E.heavy_deg;
  ea.heavy_deg_dev = &c->heavy_deg;
  auto chunk = [&]() -> int {
    for (int k = 0; k < kBfsLevelsPerSync; ++k) {
      ea.ids = list[k & 1];
      ea.out_ids = list[(k + 1) & 1];
      RG_TRY((launch_hub<BfsLevelOp, SRC_LIST>(E, op, ea)));
      RG_TRY(bfs_pull_level_dev(g, pb, sel, list[k & 1], &c->fcount, parents, &c->pheavy_count, list[(k + 1) & 1],
                                &c->ncount));
      if (checksum) hipLaunchKernelGGL(k_level_cs_ctl, ctx->num_cus, 256, 0, st, list[(k + 1) & 1], c, &c->cs);
      hipLaunchKernelGGL(k_bfs_level_end, 1, 1, 0, st, c);
      RG_HIP(hipGetLastError());
    }
    return RG_OK;
  };
  const bool captured = ctx->graphs && !ctx->timing;
  GraphCache& gc = ctx->bfs_graph;
  if (captured) {
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)g->off, (uint64_t)(uintptr_t)g->nbr, (uint64_t)(uintptr_t)g->in_off,
        (uint64_t)(uintptr_t)g->in_nbr, (uint64_t)(uintptr_t)pb.first_in, (uint64_t)pb.nn,
        (uint64_t)(uintptr_t)pb.bits, (uint64_t)(uintptr_t)pb.full[0], (uint64_t)(uintptr_t)pb.full[1],
        (uint64_t)(uintptr_t)pb.heavy, (uint64_t)pb.heavy_deg, (uint64_t)(uintptr_t)c,
        (uint64_t)(uintptr_t)parents, (uint64_t)(uintptr_t)list[0], (uint64_t)(uintptr_t)list[1],
        (uint64_t)(uintptr_t)E.heavy, (uint64_t)E.heavy_deg, (uint64_t)E.block, (uint64_t)checksum};
    if (gc.key != key || !gc.exec) {
      gc.release();
      RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = chunk();
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &graph);
      if (rc != RG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      RG_HIP(ec);
      gc.graph = graph;
      RG_HIP(hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0));
      gc.key = key;
    }
  }
  for (int64_t steps = 0;; steps += kBfsLevelsPerSync) {
    if (steps > g->n + 8) return set_error(RG_EDEVICE, "rg_bfs: no convergence");
    if (captured) RG_HIP(hipGraphLaunch(gc.exec, st));
    else RG_TRY(chunk());
    int64_t mode = 0;
    RG_TRY(ctx_read_count(ctx, &c->mode, &mode, 1));
    if (mode == BFS_DONE) break;
  }
  BfsCtl h;
  RG_TRY(ctx_copy(ctx, &h, c, sizeof(BfsCtl), hipMemcpyDeviceToHost));
  *iters = h.iter;
  *reached = h.reached;
  *edges = h.edges;
  if (cs_out) *cs_out = (uint64_t)h.cs;
  return RG_OK;
}

int bfs_push(rg_graph* g, uint32_t src, const rg_opts* o, uint32_t* parents_out, int64_t* iters_out,
             int64_t* reached_out, int64_t* edges_out, double* ms_out, uint32_t* parents_dev,
             uint64_t* checksum_out, const BfsWork* work) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if ((int64_t)src >= n) return set_error(RG_EINVAL, "rg_bfs: source out of range");
  if (work && (!parents_dev || ms_out || (o && o->variant != RG_PUSH)))
    return set_error(RG_EINVAL, "rg_bfs: a caller workspace runs the push engine into device parents");
  int64_t* counts = work ? work->counts : ctx->counts.as<int64_t>();
  Engine E;
  if (work)
    RG_TRY(make_engine(g, o, sizeof(uint32_t), E, 4, false, work->heavy, work->heavy_cap, counts + 0));
  else
    RG_TRY(make_engine(g, o, sizeof(uint32_t), E, 4));
  DevBuf& b_par = ctx->scratch[0];
  uint32_t* parents = parents_dev;
  if (!parents) {
    RG_TRY(b_par.ensure(sizeof(uint32_t) * n));
    parents = b_par.as<uint32_t>();
  }
  Frontier fa, fb;
  if (work) {
    fa.ids = work->ids[0];
    fb.ids = work->ids[1];
    fa.dense = fb.dense = work->dense;
  } else {
    RG_TRY(ctx_frontiers(ctx, n, fa, fb));
  }
  int64_t* info = counts + 4;   // [count, degsum]
  RG_TRY(fill<uint32_t>(ctx, parents, n, kNone));
  hipLaunchKernelGGL(k_set<uint32_t>, 1, 1, 0, ctx->stream, parents, (int64_t)src, src);
  hipLaunchKernelGGL(k_set<uint32_t>, 1, 1, 0, ctx->stream, fa.ids, (int64_t)0, src);
  RG_HIP(hipGetLastError());
  fa.size = 1;
  int64_t degsum = 0;
  RG_TRY(frontier_degsum(ctx, g, fa.ids, 1, &degsum, counts + 8));
  BfsOp op{parents};
  Frontier* F = &fa;
  Frontier* N = &fb;
  int64_t iter = 0;
  unsigned long long* cs = reinterpret_cast<unsigned long long*>(counts + 12);
  if (checksum_out) RG_HIP(hipMemsetAsync(cs, 0, sizeof(int64_t), ctx->stream));
  // RADAR and blocked run dense levels on the blocked layout (each unvisited
  // vertex scans its in-edges for a frontier parent; no global atomics).
  const bool blocked = o && (o->variant == RG_RADAR || o->variant == RG_BLOCKED);
  // Pull (the push-pull tree): dense levels scan the in-edges of unvisited
  // vertices with early exit -- direction-optimizing BFS.
  const bool pull = o && o->variant == RG_PULL;
  if (blocked) {
    BlockedCSC* B = nullptr;
    RG_TRY(query_layout(g, &B));   // built (once per graph) outside the timed loop
  }
  // Reached vertices and their out-degree sum accumulate from each level's
  // frontier size and degree sum (the source counts once).
  int64_t reached = 1, edges = degsum;
  std::unique_ptr<EventTimer> tm;
  if (ms_out) tm.reset(new EventTimer(ctx->stream));
  // With RG_GRAPHS_BFS_LEVELS the pull engine's levels are decided on the
  // device; otherwise the host reads every level's counts, as below.
  if (pull && ctx->bfs_device_levels && !work) {
    uint32_t* const lists[2] = {fa.ids, fb.ids};
    uint64_t c = 0;
    RG_TRY(bfs_pull_device_loop(g, E, o, src, parents, lists, checksum_out != nullptr, &iter, &reached, &edges,
                                checksum_out ? &c : nullptr));
    if (ms_out) *ms_out = tm->stop_ms();
    if (iters_out) *iters_out = iter;
    if (reached_out) *reached_out = reached;
    if (edges_out) *edges_out = edges;
    if (checksum_out) *checksum_out = c + level_hash(src, 0);
    if (parents_out) RG_TRY(rg::ctx_copy(ctx, parents_out, parents, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    return RG_OK;
  }
  // Consecutive pull levels pass the frontier on as a complete bitmap.
  const unsigned long long* pull_bits = nullptr;
  while (F->size > 0) {
    RG_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int64_t), ctx->stream));
    const unsigned long long* cur_bits = pull_bits;
    pull_bits = nullptr;
    // After a bottom-up level, stay bottom-up while that is cheaper than
    // pushing: a top-down level gathers a random parent word per frontier
    // arc (degsum of them), a bottom-up one reads every vertex's parent once
    // and scans at most the lists of the unvisited vertices (m - edges arcs
    // on a symmetric graph; an estimate on a directed one).
    const bool stay = pull && cur_bits && degsum * kPullStayArcCost > g->n + (g->m - edges);
    if (pull && degsum > 0 && (F->size + degsum > E.threshold || stay))
      RG_TRY(bfs_pull_level(g, F->ids, F->size, parents, o->heavy_degree, N->ids, info, cur_bits, &pull_bits));
    else if (blocked && degsum > 0 && F->size + degsum > E.threshold)
      RG_TRY(bfs_dense_level(g, F->ids, F->size, parents, N->ids, info));
    else
      RG_TRY(edge_map(E, op, *F, degsum, N->ids, info));
    if (checksum_out)
      hipLaunchKernelGGL(k_level_cs, ctx->num_cus, 256, 0, ctx->stream, N->ids, info, (uint32_t)(iter + 1), cs);
    int64_t h[2];
    RG_TRY(ctx_read_count(ctx, info, h, 2));
    N->size = h[0];
    degsum = h[1];
    reached += h[0];
    edges += h[1];
    std::swap(F, N);
    ++iter;
  }
  if (ms_out) *ms_out = tm->stop_ms();
  if (iters_out) *iters_out = iter;
  if (reached_out) *reached_out = reached;
  if (edges_out) *edges_out = edges;
  if (checksum_out) {
    int64_t c = 0;
    RG_TRY(ctx_read_count(ctx, reinterpret_cast<int64_t*>(cs), &c, 1));
    *checksum_out = (uint64_t)c + level_hash(src, 0);
  }
  if (parents_out) RG_TRY(rg::ctx_copy(ctx, parents_out, parents, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

// ---------------------------------------------------------------------------
// Radii (Radii.C:98-157).

__global__ void k_radii_sync(const uint32_t* ids, int64_t count, unsigned long long* vis,
                             const unsigned long long* nvis) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = ids[i];
    vis[v] = nvis[v];
  }
}

__global__ void k_radii_start(const uint32_t* starts, int k, int32_t* radii, unsigned long long* nvis) {
  // Radii.C:120-127 in start order (a repeated start keeps its last bit).
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int i = 0; i < k; ++i) {
      radii[starts[i]] = 0;
      nvis[starts[i]] = 1ull << i;
    }
}

int radii_push(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o, int32_t* radii_out,
               int64_t* rounds_out, double* ms_out);

int radii_push(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o, int32_t* radii_out,
               int64_t* rounds_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  if (nstarts < 0 || nstarts > 64) return set_error(RG_EINVAL, "rg_radii: 0..64 starts");
  for (int i = 0; i < nstarts; ++i)
    if ((int64_t)starts[i] >= n) return set_error(RG_EINVAL, "rg_radii: start out of range");
  Engine E;
  RG_TRY(make_engine(g, o, sizeof(uint64_t), E, 4));
  DevBuf &b_rad = ctx->scratch[0], &b_vis = ctx->scratch[1], &b_nvis = ctx->scratch[2];
  RG_TRY(b_rad.ensure(sizeof(int32_t) * n));
  RG_TRY(b_vis.ensure(sizeof(uint64_t) * n));
  RG_TRY(b_nvis.ensure(sizeof(uint64_t) * n));
  int32_t* radii = b_rad.as<int32_t>();
  auto* vis = b_vis.as<unsigned long long>();
  auto* nvis = b_nvis.as<unsigned long long>();
  RG_TRY(fill<int32_t>(ctx, radii, n, -1));
  RG_HIP(hipMemsetAsync(vis, 0, sizeof(uint64_t) * n, ctx->stream));
  RG_HIP(hipMemsetAsync(nvis, 0, sizeof(uint64_t) * n, ctx->stream));
  Frontier fa, fb;
  RG_TRY(ctx_frontiers(ctx, n, fa, fb));
  if (nstarts > 0) {
    RG_TRY(rg::ctx_copy(ctx, fa.ids, starts, sizeof(uint32_t) * nstarts, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_radii_start, 1, 64, 0, ctx->stream, fa.ids, nstarts, radii, nvis);
  fa.size = nstarts;
  int64_t degsum = 0;
  RG_TRY(frontier_degsum(ctx, g, fa.ids, nstarts, &degsum));
  int64_t* info = ctx->counts.as<int64_t>() + 4;
  Frontier* F = &fa;
  Frontier* N = &fb;
  int32_t round = 0;
  EventTimer tm(ctx->stream);
  while (F->size > 0) {
    round++;
    hipLaunchKernelGGL(k_radii_sync, grid_for(ctx, F->size), 256, 0, ctx->stream, F->ids, F->size, vis,
                       nvis);
    RG_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int64_t), ctx->stream));
    RadiiOp op{vis, nvis, radii, round};
    RG_TRY(edge_map(E, op, *F, degsum, N->ids, info));
    int64_t h[2];
    RG_TRY(ctx_read_count(ctx, info, h, 2));
    N->size = h[0];
    degsum = h[1];
    std::swap(F, N);
  }
  double ms = tm.stop_ms();
  if (rounds_out) *rounds_out = round;
  if (ms_out) *ms_out = ms;
  if (radii_out) RG_TRY(rg::ctx_copy(ctx, radii_out, radii, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

// ---------------------------------------------------------------------------
// BC (BC-manyiter-v2.C:124-241).

__global__ void k_bc_start(double* np, uint8_t* visited, uint32_t start) {
  np[start] = 1.0;
  visited[start] = 1;
}

__global__ void k_mark(const uint32_t* ids, int64_t count, uint8_t* visited) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    visited[ids[i]] = 1;
}

// BC_Back_Vertex_F (BC-manyiter-v2.C:93-101).
__global__ void k_bc_back_vertex(const uint32_t* ids, int64_t count, uint8_t* visited, double* deps,
                                 const double* inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = ids[i];
    visited[v] = 1;
    deps[v] += inv[v];
  }
}

__global__ void k_invert(double* np, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    np[i] = 1 / np[i];
}

__global__ void k_bc_final(double* deps, const double* inv, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    deps[i] = (deps[i] - inv[i]) / inv[i];
}

int bc_push(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
            int64_t* combined_out, double* ms_out);

int bc_push(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
            int64_t* combined_out, double* ms_out) {
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  for (int i = 0; i < nsources; ++i)
    if ((int64_t)sources[i] >= n) return set_error(RG_EINVAL, "rg_bc: source out of range");
  DevBuf &b_np = ctx->scratch[0], &b_dep = ctx->scratch[1], &b_vis = ctx->scratch[2],
         &b_lv = ctx->scratch[3];
  RG_TRY(b_np.ensure(sizeof(double) * n));
  RG_TRY(b_dep.ensure(sizeof(double) * n));
  RG_TRY(b_vis.ensure((size_t)n));
  RG_TRY(b_lv.ensure(sizeof(uint32_t) * (n + 1)));
  double* np = b_np.as<double>();
  double* deps = b_dep.as<double>();
  uint8_t* visited = b_vis.as<uint8_t>();
  uint32_t* lv = b_lv.as<uint32_t>();
  Frontier F, unused;   // F views the level lists stored in lv
  RG_TRY(ctx_frontiers(ctx, n, F, unused));
  RG_HIP(hipMemsetAsync(deps, 0, sizeof(double) * n, ctx->stream));
  int64_t* info = ctx->counts.as<int64_t>() + 4;
  int64_t combined = 0;
  const bool was_transposed = g->transposed;
  // RADAR and blocked run dense levels on the blocked layout (bcdense.hip);
  // push and HUBDUP keep the atomic push engine for every level.
  const bool blocked = o && (o->variant == RG_RADAR || o->variant == RG_BLOCKED);
  // Blocked dense levels keep one complete frontier bitmap per level (the
  // forward's dense levels write the next one themselves, sparse levels set
  // it from their list) and the backward reuses them: no level rebuilds a
  // bitmap from a list.  Levels past kLevelBits fall back to doing so.
  constexpr int64_t kLevelBits = 64;
  const int64_t nwords = n / 64 + 1;
  DevBuf& b_lvb = ctx->scratch[4];
  unsigned long long* lvb = nullptr;
  if (blocked) {
    BlockedCSC* B = nullptr;
    RG_TRY(query_layout(g, &B));   // the forward orientation's layout, outside the timed loop
    RG_TRY(b_lvb.ensure(sizeof(unsigned long long) * nwords * kLevelBits));
    lvb = b_lvb.as<unsigned long long>();
  }
  auto level_bits = [&](int64_t L) -> unsigned long long* {
    return lvb && L < kLevelBits ? lvb + L * nwords : nullptr;
  };
  EventTimer tm(ctx->stream);
  for (int it = 0; it < nsources; ++it) {
    Engine E;
    RG_TRY(make_engine(g, o, sizeof(double), E, 4));
    uint32_t start = sources[it];
    RG_HIP(hipMemsetAsync(np, 0, sizeof(double) * n, ctx->stream));
    RG_HIP(hipMemsetAsync(visited, 0, (size_t)n, ctx->stream));
    hipLaunchKernelGGL(k_bc_start, 1, 1, 0, ctx->stream, np, visited, start);
    hipLaunchKernelGGL(k_set<uint32_t>, 1, 1, 0, ctx->stream, lv, (int64_t)0, start);
    std::vector<int64_t> lstart = {0, 1};
    int64_t degsum = 0;
    RG_TRY(frontier_degsum(ctx, g, lv, 1, &degsum));
    std::vector<int64_t> lvdeg = {degsum};   // out-degree sum of each level
    if (unsigned long long* b0 = level_bits(0)) {
      RG_TRY(fill<unsigned long long>(ctx, b0, nwords, 0ull));
      RG_TRY(set_frontier_bits(ctx, lv, 1, b0));
    }
    int64_t round = 0;
    BcFwdOp fop{np, visited};
    while (true) {
      int64_t fb = lstart[lstart.size() - 2], fe = lstart.back();
      if (fe == fb) break;
      round++;
      F.ids = lv + fb;
      F.size = fe - fb;
      RG_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int64_t), ctx->stream));
      // Level round - 1 is the frontier; level round is being found.
      unsigned long long* cur_b = level_bits(round - 1);
      unsigned long long* nxt_b = cur_b ? level_bits(round) : nullptr;
      if (nxt_b) RG_TRY(fill<unsigned long long>(ctx, nxt_b, nwords, 0ull));
      const bool dense = blocked && F.size + degsum > E.threshold;
      if (dense)
        RG_TRY(bc_dense_level(g, true, F.ids, F.size, np, np, visited, lv + fe, info, cur_b, nxt_b));
      else
        RG_TRY(edge_map(E, fop, F, degsum, lv + fe, info));
      int64_t h[2];
      RG_TRY(ctx_read_count(ctx, info, h, 2));
      if (h[0] > 0)
        hipLaunchKernelGGL(k_mark, grid_for(ctx, h[0]), 256, 0, ctx->stream, lv + fe, h[0], visited);
      if (!dense && nxt_b) RG_TRY(set_frontier_bits(ctx, lv + fe, h[0], nxt_b));
      degsum = h[1];
      lvdeg.push_back(degsum);
      lstart.push_back(fe + h[0]);
    }
    combined += round;
    hipLaunchKernelGGL(k_invert, grid_for(ctx, n), 256, 0, ctx->stream, np, n);
    RG_HIP(hipMemsetAsync(visited, 0, (size_t)n, ctx->stream));
    // Frontier = Levels[round-1]; level r is lv[lstart[r], lstart[r+1]).
    auto level = [&](int64_t r) { F.ids = lv + lstart[r]; F.size = lstart[r + 1] - lstart[r]; };
    level(round - 1);
    if (F.size > 0)
      hipLaunchKernelGGL(k_bc_back_vertex, grid_for(ctx, F.size), 256, 0, ctx->stream, F.ids, F.size,
                         visited, deps, np);
    RG_TRY(rg_graph_transpose(g));
    Engine EB;
    RG_TRY(make_engine(g, o, sizeof(double), EB, 4, true));
    BcBackOp bop{deps, visited};
    for (int64_t r = round - 2; r >= 0; --r) {
      // The frontier is level r + 1; its out-degree sum in the transposed
      // graph is its in-degree sum, the forward's own sum on a symmetric graph.
      int64_t ds = 0;
      if (g->symmetric) ds = lvdeg[r + 1];
      else RG_TRY(frontier_degsum(ctx, g, F.ids, F.size, &ds));
      if (blocked && ds > 0 && F.size + ds > EB.threshold)
        RG_TRY(bc_dense_level(g, false, F.ids, F.size, deps, deps, visited, nullptr, nullptr, level_bits(r + 1)));
      else
        RG_TRY(edge_map(EB, bop, F, ds, nullptr, nullptr));
      level(r);
      hipLaunchKernelGGL(k_bc_back_vertex, grid_for(ctx, F.size), 256, 0, ctx->stream, F.ids, F.size,
                         visited, deps, np);
    }
    hipLaunchKernelGGL(k_bc_final, grid_for(ctx, n), 256, 0, ctx->stream, deps, np, n);
    RG_HIP(hipGetLastError());
  }
  double ms = tm.stop_ms();
  (void)was_transposed;
  if (combined_out) *combined_out = combined;
  if (ms_out) *ms_out = ms;
  ctx->last_result = deps;
  if (deps_out) RG_TRY(rg::ctx_copy(ctx, deps_out, deps, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

// ---------------------------------------------------------------------------
// Pull PageRank of the push-pull tree: threshold 0 makes every iteration an
// edgeMapDense (push-pull Ligra/apps/PageRank.C:122, ligra.h:61-88), which
// for each vertex v sums p_curr[s]/outdeg(s) over its in-neighbours with the
// non-atomic update.  Light vertices (in-degree <= 64) sum sequentially in
// in-list order -- the reference's order -- one thread each; heavier ones are
// reduced by a wavefront each.

constexpr uint32_t kPullLight = 64;

__global__ void k_pr_pull_light(const uint64_t* in_off, const uint32_t* in_nbr, int64_t n,
                                const double* contrib, double* acc, uint32_t* heavy,
                                int64_t* heavy_count, const int64_t* done) {
  if (*done) return;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t b = in_off[v], e = in_off[v + 1];
    if (e - b > kPullLight) {
      int64_t slot = atomicAdd((unsigned long long*)heavy_count, 1ull);
      heavy[slot] = (uint32_t)v;
      continue;
    }
    double s = acc[v];
    for (uint64_t j = b; j < e; ++j) s += contrib[in_nbr[j]];
    acc[v] = s;
  }
}

__global__ void k_pr_pull_heavy(const uint64_t* in_off, const uint32_t* in_nbr, const double* contrib,
                                double* acc, const uint32_t* heavy, const int64_t* heavy_count,
                                const int64_t* done) {
  if (*done) return;
  const int lane = (int)__lane_id();
  const int64_t items = *heavy_count;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t it = wid; it < items; it += nw) {
    uint32_t v = heavy[it];
    double s = 0.0;
    for (uint64_t j = in_off[v] + lane; j < in_off[v + 1]; j += kWave) s += contrib[in_nbr[j]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) acc[v] += s;
  }
}

int pagerank_pull(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                  int64_t* iters_out, double* l1_out, double* ms_out) {
  (void)o;
  rg_ctx* ctx = g->ctx;
  const int64_t n = g->n;
  DevBuf &b_rank = ctx->scratch[0], &b_contrib = ctx->scratch[1], &b_acc = ctx->scratch[2],
         &b_part = ctx->scratch[3], &b_heavy = ctx->scratch[4];
  RG_TRY(b_rank.ensure(sizeof(double) * n));
  RG_TRY(b_contrib.ensure(sizeof(double) * n));
  RG_TRY(b_acc.ensure(sizeof(double) * n));
  RG_TRY(b_heavy.ensure(sizeof(uint32_t) * n));
  const int vgrid = std::min<int>(grid_for(ctx, n), 1024);
  RG_TRY(b_part.ensure(sizeof(double) * (vgrid + std::max<int64_t>(maxiters, 1) + 8)));
  double* rank = b_rank.as<double>();
  double* contrib = b_contrib.as<double>();
  double* acc = b_acc.as<double>();
  double* partial = b_part.as<double>();
  double* l1_hist = partial + vgrid;
  int64_t* done = ctx->counts.as<int64_t>() + 16;
  int64_t* hcount = ctx->counts.as<int64_t>() + 17;
  RG_HIP(hipMemsetAsync(done, 0, sizeof(int64_t), ctx->stream));
  const double damping = 0.85, epsilon = 0.0000001;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * (1 / (double)n);
  hipLaunchKernelGGL(k_pr_init, grid_for(ctx, n), 256, 0, ctx->stream, g->off, n, rank, contrib, acc,
                     one_over_n);
  RG_HIP(hipGetLastError());
  EventTimer tm(ctx->stream);
  for (int64_t it = 1; it <= maxiters; ++it) {
    RG_HIP(hipMemsetAsync(hcount, 0, sizeof(int64_t), ctx->stream));
    hipLaunchKernelGGL(k_pr_pull_light, grid_for(ctx, n), 256, 0, ctx->stream, g->in_off, g->in_nbr, n,
                       contrib, acc, b_heavy.as<uint32_t>(), hcount, done);
    hipLaunchKernelGGL(k_pr_pull_heavy, ctx->num_cus * 8, 256, 0, ctx->stream, g->in_off, g->in_nbr,
                       contrib, acc, b_heavy.as<uint32_t>(), hcount, done);
    hipLaunchKernelGGL(k_pr_vertex, vgrid, 256, 0, ctx->stream, g->off, n, rank, contrib, acc, damping,
                       added, partial, done);
    hipLaunchKernelGGL(k_pr_finish, 1, 256, 0, ctx->stream, partial, vgrid, l1_hist, it, done, epsilon);
    RG_HIP(hipGetLastError());
  }
  double ms = tm.stop_ms();
  int64_t dn = 0;
  RG_TRY(ctx_read_count(ctx, done, &dn, 1));
  int64_t last = dn ? dn : maxiters;
  double l1 = 0.0;
  if (last >= 1) RG_TRY(rg::ctx_copy(ctx, &l1, l1_hist + (last - 1), sizeof(double), hipMemcpyDeviceToHost));
  if (iters_out) *iters_out = dn ? dn : maxiters + 1;
  if (l1_out) *l1_out = l1;
  if (ms_out) *ms_out = ms;
  ctx->last_result = rank;
  if (ranks_out) RG_TRY(rg::ctx_copy(ctx, ranks_out, rank, sizeof(double) * n, hipMemcpyDeviceToHost));
  return RG_OK;
}

}  // namespace rg

using rg::set_error;

extern "C" {

void rg_opts_default(rg_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->variant = RG_PUSH;
  o->hub_placement = RG_HUBS_LDS;
  o->hubs = 0;
  o->threshold = -1;
  o->wg_size = 0;
  o->heavy_degree = 0;
}

int rg_pagerank(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out, int64_t* iters_out,
                double* l1_out, double* ms_out) {
  if (!g || maxiters < 0) return set_error(RG_EINVAL, "rg_pagerank: bad arguments");
  RG_HIP(hipSetDevice(g->ctx->device));
  if (o && o->variant == RG_PULL)
    return rg::pagerank_pull(g, maxiters, o, ranks_out, iters_out, l1_out, ms_out);
  if (o && o->variant == RG_BLOCKED)
    return rg::pagerank_blocked(g, maxiters, o, ranks_out, iters_out, l1_out, ms_out);
  return rg::pagerank_push(g, maxiters, o, ranks_out, iters_out, l1_out, ms_out);
}

int rg_bfs(rg_graph* g, uint32_t src, const rg_opts* o, uint32_t* parents_out, int64_t* iters_out,
           int64_t* reached_out, int64_t* edges_out, double* ms_out) {
  if (!g) return set_error(RG_EINVAL, "rg_bfs: null graph");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::bfs_push(g, src, o, parents_out, iters_out, reached_out, edges_out, ms_out, nullptr, nullptr);
}

int rg_bfs_query(rg_graph* g, uint32_t src, const rg_opts* o, rg_bfs_summary* out, double* ms_out) {
  if (!g || !out) return rg::set_error(RG_EINVAL, "rg_bfs_query: null argument");
  RG_HIP(hipSetDevice(g->ctx->device));
  uint64_t cs = 0;
  RG_TRY(rg::bfs_push(g, src, o, nullptr, &out->iters, &out->reached, &out->edges, ms_out, nullptr, &cs));
  out->checksum = cs;
  return RG_OK;
}

int rg_pagerank_delta(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out, int64_t* iters_out,
                      double* l1_out, double* ms_out) {
  if (!g || maxiters < 0) return set_error(RG_EINVAL, "rg_pagerank_delta: bad arguments");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::pagerank_delta(g, maxiters, o, ranks_out, iters_out, l1_out, ms_out);
}

int rg_radii(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o, int32_t* radii_out,
             int64_t* rounds_out, double* ms_out) {
  if (!g || (nstarts > 0 && !starts)) return set_error(RG_EINVAL, "rg_radii: bad arguments");
  RG_HIP(hipSetDevice(g->ctx->device));
  if (o && (o->variant == RG_RADAR || o->variant == RG_BLOCKED))
    return rg::radii_batch(g, starts, nstarts, o, radii_out, rounds_out, ms_out);
  return rg::radii_push(g, starts, nstarts, o, radii_out, rounds_out, ms_out);
}

int rg_bc(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
          int64_t* combined_out, double* ms_out) {
  if (!g || nsources < 0 || (nsources > 0 && !sources)) return set_error(RG_EINVAL, "rg_bc: bad arguments");
  RG_HIP(hipSetDevice(g->ctx->device));
  return rg::bc_push(g, sources, nsources, o, deps_out, combined_out, ms_out);
}

}  // extern "C"

// ===========================================================================
// Engine-level C ABI: Ligra's vertexSubset / edgeMap / vertexMap with the
// reference's functors as named operators (PNO Ligra/ligra/ligra.h:336-340,
// 447-479; vertexSubset.h:113-229).

struct rg_subset {
  rg_graph* g = nullptr;
  rg::DevBuf ids;
  int64_t size = 0;
  int64_t degsum = -1;   // sum of out-degrees, -1 = unknown
  bool all = false;
};

struct rg_array {
  rg_ctx* ctx = nullptr;
  rg::DevBuf buf;
  int64_t count = 0;
  int elem = 0;
};

namespace rg {

__global__ void k_bc_back_vertex_all(int64_t n, uint8_t* visited, double* deps, const double* inv) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    visited[v] = 1;
    deps[v] += inv[v];
  }
}

__global__ void k_pr_vertex_update(const uint32_t* ids, int64_t count, bool all, double* p_next, double damping,
                                   double added) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    p_next[v] = __dadd_rn(__dmul_rn(damping, p_next[v]), added);
  }
}

__global__ void k_sync_u64(const uint32_t* ids, int64_t count, bool all, unsigned long long* dst,
                           const unsigned long long* src) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    dst[v] = src[v];
  }
}

__global__ void k_mark_u8(const uint32_t* ids, int64_t count, bool all, uint8_t* dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[all ? i : ids[i]] = 1;
}

__global__ void k_bc_back_vertex_l(const uint32_t* ids, int64_t count, uint8_t* visited, double* deps,
                                   const double* inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = ids[i];
    visited[v] = 1;
    deps[v] += inv[v];
  }
}

__global__ void k_pr_contrib(const uint64_t* off, int64_t n, const double* p, double* contrib) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t deg = off[v + 1] - off[v];
    contrib[v] = deg ? p[v] / (double)(uint32_t)deg : 0.0;
  }
}

// x = |x - y| (PageRank.C:125-127) and x = |y| (PageRankDelta.C:136-137).
__global__ void k_abs_diff(const uint32_t* ids, int64_t count, bool all, double* x, const double* y, int diff) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    x[v] = diff ? fabs(x[v] - y[v]) : fabs(y[v]);
  }
}

// ---- vertexFilter (ligra.h:500-509) ---------------------------------------
// Filter functors write a keep flag per vertex of the input subset; the
// flags are then compacted into an ascending id list (the order toSparse
// gives Ligra's dense output).

// PR_Vertex_F_FirstRound (PageRankDelta.C:49-63) and PR_Vertex_F (65-77);
// multiply and add rounded separately, as the reference's host code.
__global__ void k_filter_prd(const uint32_t* ids, int64_t count, bool all, double* p, double* delta,
                             const double* ngh, double damping, double added, double one_over_n,
                             double eps2, int first, uint8_t* keep) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    bool k;
    if (first) {
      double d = __dadd_rn(__dmul_rn(damping, ngh[v]), added);
      double pv = p[v] + d;
      p[v] = pv;
      d = d - one_over_n;
      delta[v] = d;
      k = fabs(d) > __dmul_rn(eps2, pv);
    } else {
      double d = __dmul_rn(ngh[v], damping);
      delta[v] = d;
      k = fabs(d) > __dmul_rn(eps2, p[v]);
      if (k) p[v] += d;
    }
    keep[v] = k;
  }
}

// Keeps v with x[v] == value (4-byte elements) or x[v] != 0 (1-byte).
__global__ void k_filter_u32_eq(const uint32_t* ids, int64_t count, bool all, const uint32_t* x, uint32_t value,
                                uint8_t* keep) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    keep[v] = x[v] == value;
  }
}
__global__ void k_filter_u8_nz(const uint32_t* ids, int64_t count, bool all, const uint8_t* x, uint8_t* keep) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = all ? (uint32_t)i : ids[i];
    keep[v] = x[v] != 0;
  }
}

constexpr int kCompactBlock = 256;
constexpr int kCompactBlocks = 1024;

// Pass 1: kept flags per contiguous chunk of vertices (one workgroup each).
__global__ __launch_bounds__(kCompactBlock) void k_compact_count(const uint8_t* keep, int64_t n, int64_t chunk,
                                                                 int64_t* counts) {
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  int64_t c = 0;
  for (int64_t v = lo + threadIdx.x; v < hi; v += kCompactBlock) c += keep[v] != 0;
  __shared__ int64_t red[kCompactBlock];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = kCompactBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// Pass 2 (one workgroup): exclusive prefix of the chunk counts in place,
// the total in counts[nb].
__global__ __launch_bounds__(kCompactBlocks) void k_compact_scan(int64_t* counts, int nb) {
  __shared__ int64_t s[kCompactBlocks];
  const int t = threadIdx.x;
  int64_t x = t < nb ? counts[t] : 0;
  s[t] = x;
  __syncthreads();
  for (int d = 1; d < kCompactBlocks; d <<= 1) {
    int64_t y = t >= d ? s[t - d] : 0;
    __syncthreads();
    s[t] += y;
    __syncthreads();
  }
  if (t < nb) counts[t] = s[t] - x;
  if (t == kCompactBlocks - 1) counts[nb] = s[t];
}

// Pass 3: each workgroup writes its chunk's kept ids in ascending order and
// clears the flags it read (the flag array is all zero between calls).
__global__ __launch_bounds__(kCompactBlock) void k_compact_write(uint8_t* keep, int64_t n, int64_t chunk,
                                                                 const int64_t* offsets, uint32_t* out) {
  constexpr int kW = kCompactBlock / kWave;
  __shared__ uint32_t wsum[kW];
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  int64_t pos = offsets[blockIdx.x];
  const int lane = (int)__lane_id(), w = threadIdx.x / kWave;
  for (int64_t t0 = lo; t0 < hi; t0 += kCompactBlock) {
    const int64_t v = t0 + threadIdx.x;
    const bool k = v < hi && keep[v];
    if (k) keep[v] = 0;
    const unsigned long long bal = __ballot(k);
    const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < kW; ++i) {
      before += i < w ? wsum[i] : 0u;
      total += wsum[i];
    }
    if (k) out[pos + before + below] = (uint32_t)v;
    pos += total;
    __syncthreads();
  }
}

// Deterministic FP64 sum (sequence::plusReduce): fixed per-workgroup
// partials, then one workgroup adds them in a fixed order.
__global__ __launch_bounds__(256) void k_sum_partial(const double* x, int64_t n, double* part) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_sum_final(double* part, int np) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[np] = red[0];
}

static int check_state(rg_graph* g, rg_array* const* st, int nstate, std::initializer_list<int> elems,
                       const char* who) {
  if (nstate != (int)elems.size()) return set_error(RG_EINVAL, std::string(who) + ": wrong number of state arrays");
  int i = 0;
  for (int e : elems) {
    if (!st[i] || st[i]->elem != e || st[i]->count < g->n)
      return set_error(RG_EINVAL, std::string(who) + ": state array " + std::to_string(i) +
                                      " must hold n elements of " + std::to_string(e) + " bytes");
    ++i;
  }
  return RG_OK;
}

}  // namespace rg

extern "C" {

int rg_subset_create(rg_graph* g, rg_subset** out) {
  if (!g || !out) return set_error(RG_EINVAL, "rg_subset_create: null argument");
  std::unique_ptr<rg_subset> s(new rg_subset);
  s->g = g;
  RG_TRY(s->ids.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(g->n, 64)));
  *out = s.release();
  return RG_OK;
}

int rg_subset_set(rg_subset* s, const uint32_t* ids, int64_t count) {
  if (!s || count < 0 || count > s->g->n || (count > 0 && !ids)) return set_error(RG_EINVAL, "rg_subset_set: bad arguments");
  for (int64_t i = 0; i < count; ++i)
    if ((int64_t)ids[i] >= s->g->n) return set_error(RG_EINVAL, "rg_subset_set: vertex out of range");
  if (count) RG_TRY(rg::ctx_copy(s->g->ctx, s->ids.p, ids, sizeof(uint32_t) * count, hipMemcpyHostToDevice));
  s->size = count;
  s->all = false;
  s->degsum = -1;
  return RG_OK;
}

int rg_subset_set_all(rg_subset* s) {
  if (!s) return set_error(RG_EINVAL, "rg_subset_set_all: null");
  s->all = true;
  s->size = s->g->n;
  s->degsum = s->g->m;
  return RG_OK;
}

int rg_subset_size(const rg_subset* s, int64_t* size) {
  if (!s || !size) return set_error(RG_EINVAL, "rg_subset_size: null");
  *size = s->size;
  return RG_OK;
}

int rg_subset_get(const rg_subset* s, uint32_t* ids) {
  if (!s || !ids) return set_error(RG_EINVAL, "rg_subset_get: null");
  if (s->all) {
    for (int64_t i = 0; i < s->size; ++i) ids[i] = (uint32_t)i;
  } else if (s->size) {
    RG_TRY(rg::ctx_copy(s->g->ctx, ids, s->ids.p, sizeof(uint32_t) * s->size, hipMemcpyDeviceToHost));
  }
  return RG_OK;
}

int rg_subset_free(rg_subset* s) {
  delete s;
  return RG_OK;
}

int rg_array_alloc(rg_ctx* ctx, int64_t count, int elem_bytes, rg_array** out) {
  if (!ctx || !out || count < 0 || (elem_bytes != 1 && elem_bytes != 4 && elem_bytes != 8))
    return set_error(RG_EINVAL, "rg_array_alloc: bad arguments");
  std::unique_ptr<rg_array> a(new rg_array);
  a->ctx = ctx;
  a->count = count;
  a->elem = elem_bytes;
  RG_TRY(a->buf.ensure((size_t)std::max<int64_t>(count, 1) * elem_bytes));
  *out = a.release();
  return RG_OK;
}

int rg_array_write(rg_array* a, const void* host, int64_t first, int64_t count) {
  if (!a || (count > 0 && !host) || first < 0 || first + count > a->count)
    return set_error(RG_EINVAL, "rg_array_write: bad range");
  if (count)
    RG_TRY(rg::ctx_copy(a->ctx, static_cast<char*>(a->buf.p) + first * a->elem, host, (size_t)count * a->elem,
                     hipMemcpyHostToDevice));
  return RG_OK;
}

int rg_array_read(const rg_array* a, void* host, int64_t first, int64_t count) {
  if (!a || (count > 0 && !host) || first < 0 || first + count > a->count)
    return set_error(RG_EINVAL, "rg_array_read: bad range");
  RG_HIP(hipStreamSynchronize(a->ctx->stream));
  if (count)
    RG_TRY(rg::ctx_copy(a->ctx, host, static_cast<const char*>(a->buf.p) + first * a->elem, (size_t)count * a->elem,
                     hipMemcpyDeviceToHost));
  return RG_OK;
}

int rg_array_fill(rg_array* a, uint64_t bits) {
  if (!a) return set_error(RG_EINVAL, "rg_array_fill: null");
  rg_ctx* ctx = a->ctx;
  switch (a->elem) {
    case 1: return rg::fill<uint8_t>(ctx, a->buf.as<uint8_t>(), a->count, (uint8_t)bits);
    case 4: return rg::fill<uint32_t>(ctx, a->buf.as<uint32_t>(), a->count, (uint32_t)bits);
    default: return rg::fill<unsigned long long>(ctx, a->buf.as<unsigned long long>(), a->count, bits);
  }
}

int rg_array_free(rg_array* a) {
  delete a;
  return RG_OK;
}

// edgeMap (ligra.h:264-302, 336-340).  Operators and their state arrays:
//   RG_OP_BFS      [parents u32]                       BFS_F (BFS.C:28-49)
//   RG_OP_PAGERANK [contrib f64, p_next f64]           PR_F (PageRank.C:30-43); contrib[s] =
//                                                      p_curr[s]/outdeg(s) (rg_vertexmap RG_VOP_PR_CONTRIB)
//   RG_OP_RADII    [visited u64, next_visited u64, radii i32]   Radii_F (Radii.C:35-64), `round`
//   RG_OP_BC_FWD   [num_paths f64, visited u8]         BC_F (BC-manyiter-v2.C:32-56)
//   RG_OP_BC_BACK  [dependencies f64, visited u8]      BC_Back_F (BC-manyiter-v2.C:58-79)
int rg_edgemap(rg_graph* g, int op, rg_subset* in, rg_array* const* state, int nstate, int64_t round,
               int64_t threshold, int flags, const rg_opts* o, rg_subset* out) {
  if (!g || !in || in->g->n != g->n) return set_error(RG_EINVAL, "edgeMap: Sizes Don't match");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  rg_opts oo;
  if (o) oo = *o;
  else rg_opts_default(&oo);
  oo.threshold = threshold;
  if (oo.variant == RG_PULL || oo.variant == RG_BLOCKED) oo.variant = RG_PUSH;
  rg::Frontier F, unused;
  RG_TRY(rg::ctx_frontiers(ctx, g->n, F, unused));
  F.ids = in->ids.as<uint32_t>();
  F.size = in->size;
  F.all = in->all;
  int64_t degsum = in->degsum;
  if (!in->all && degsum < 0) RG_TRY(rg::frontier_degsum(ctx, g, F.ids, F.size, &degsum));
  const bool want_out = out && !(flags & RG_NO_OUTPUT);
  if (want_out && out->g->n != g->n) return set_error(RG_EINVAL, "edgeMap: Sizes Don't match");
  int64_t* info = ctx->counts.as<int64_t>() + 28;
  RG_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int64_t), ctx->stream));
  uint32_t* out_ids = want_out ? out->ids.as<uint32_t>() : ctx->fr_ids[1].as<uint32_t>();
  rg::Engine E;
  switch (op) {
    case RG_OP_BFS: {
      RG_TRY(rg::check_state(g, state, nstate, {4}, "edgeMap(BFS_F)"));
      RG_TRY(rg::make_engine(g, &oo, sizeof(uint32_t), E, 4));
      RG_TRY(rg::edge_map(E, rg::BfsOp{state[0]->buf.as<uint32_t>()}, F, degsum, out_ids, info));
      break;
    }
    case RG_OP_PAGERANK: {
      RG_TRY(rg::check_state(g, state, nstate, {8, 8}, "edgeMap(PR_F)"));
      RG_TRY(rg::make_engine(g, &oo, sizeof(double), E, 1, true));
      int64_t* zero = ctx->counts.as<int64_t>() + 30;
      RG_HIP(hipMemsetAsync(zero, 0, sizeof(int64_t), ctx->stream));
      rg::PrOp op2{state[0]->buf.as<double>(), state[1]->buf.as<double>(), zero};
      RG_TRY(rg::edge_map(E, op2, F, degsum, nullptr, nullptr));
      break;
    }
    case RG_OP_RADII: {
      RG_TRY(rg::check_state(g, state, nstate, {8, 8, 4}, "edgeMap(Radii_F)"));
      RG_TRY(rg::make_engine(g, &oo, sizeof(uint64_t), E, 4));
      rg::RadiiOp op2{state[0]->buf.as<unsigned long long>(), state[1]->buf.as<unsigned long long>(),
                      state[2]->buf.as<int32_t>(), (int32_t)round};
      RG_TRY(rg::edge_map(E, op2, F, degsum, out_ids, info));
      break;
    }
    case RG_OP_BC_FWD: {
      RG_TRY(rg::check_state(g, state, nstate, {8, 1}, "edgeMap(BC_F)"));
      RG_TRY(rg::make_engine(g, &oo, sizeof(double), E, 4));
      RG_TRY(rg::edge_map(E, rg::BcFwdOp{state[0]->buf.as<double>(), state[1]->buf.as<uint8_t>()}, F, degsum,
                          out_ids, info));
      break;
    }
    case RG_OP_BC_BACK: {
      RG_TRY(rg::check_state(g, state, nstate, {8, 1}, "edgeMap(BC_Back_F)"));
      RG_TRY(rg::make_engine(g, &oo, sizeof(double), E, 4, true));
      RG_TRY(rg::edge_map(E, rg::BcBackOp{state[0]->buf.as<double>(), state[1]->buf.as<uint8_t>()}, F, degsum,
                          nullptr, nullptr));
      break;
    }
    default:
      return set_error(RG_EINVAL, "edgeMap: unknown operator");
  }
  if (want_out) {
    int64_t h[2] = {0, 0};
    if (op == RG_OP_BFS || op == RG_OP_RADII || op == RG_OP_BC_FWD) RG_TRY(rg::ctx_read_count(ctx, info, h, 2));
    out->size = h[0];
    out->degsum = h[1];
    out->all = false;
  } else if (out) {
    out->size = 0;
    out->degsum = 0;
    out->all = false;
  }
  return RG_OK;
}

// vertexMap (ligra.h:464-479) with the reference's vertex functors.
int rg_vertexmap(rg_graph* g, int vop, rg_subset* s, rg_array* const* state, int nstate, double a, double b) {
  if (!g || !s || s->g->n != g->n) return set_error(RG_EINVAL, "vertexMap: bad subset");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  const int64_t cnt = s->all ? g->n : s->size;
  if (cnt == 0) return RG_OK;
  const int grid = rg::grid_for(ctx, cnt);
  const uint32_t* ids = s->ids.as<uint32_t>();
  switch (vop) {
    case RG_VOP_PR_UPDATE:   // PR_Vertex_F (PageRank.C:46-58): p = a*p + b
      RG_TRY(rg::check_state(g, state, nstate, {8}, "vertexMap(PR_Vertex_F)"));
      hipLaunchKernelGGL(rg::k_pr_vertex_update, grid, 256, 0, ctx->stream, ids, cnt, s->all,
                         state[0]->buf.as<double>(), a, b);
      break;
    case RG_VOP_PR_CONTRIB:  // contrib[v] = p[v] / outdeg(v), all vertices
      RG_TRY(rg::check_state(g, state, nstate, {8, 8}, "vertexMap(contrib)"));
      hipLaunchKernelGGL(rg::k_pr_contrib, rg::grid_for(ctx, g->n), 256, 0, ctx->stream, g->off, g->n,
                         state[0]->buf.as<double>(), state[1]->buf.as<double>());
      break;
    case RG_VOP_RADII_SYNC:  // Radii_Vertex_F (Radii.C:67-75): Visited = NextVisited
      RG_TRY(rg::check_state(g, state, nstate, {8, 8}, "vertexMap(Radii_Vertex_F)"));
      hipLaunchKernelGGL(rg::k_sync_u64, grid, 256, 0, ctx->stream, ids, cnt, s->all,
                         state[0]->buf.as<unsigned long long>(), state[1]->buf.as<unsigned long long>());
      break;
    case RG_VOP_MARK:        // BC_Vertex_F (BC-manyiter-v2.C:82-89): Visited = 1
      RG_TRY(rg::check_state(g, state, nstate, {1}, "vertexMap(BC_Vertex_F)"));
      hipLaunchKernelGGL(rg::k_mark_u8, grid, 256, 0, ctx->stream, ids, cnt, s->all, state[0]->buf.as<uint8_t>());
      break;
    case RG_VOP_BC_BACK:     // BC_Back_Vertex_F (BC-manyiter-v2.C:93-101)
      RG_TRY(rg::check_state(g, state, nstate, {1, 8, 8}, "vertexMap(BC_Back_Vertex_F)"));
      if (s->all)
        hipLaunchKernelGGL(rg::k_bc_back_vertex_all, grid, 256, 0, ctx->stream, g->n, state[0]->buf.as<uint8_t>(),
                           state[1]->buf.as<double>(), state[2]->buf.as<double>());
      else
        hipLaunchKernelGGL(rg::k_bc_back_vertex_l, grid, 256, 0, ctx->stream, ids, cnt, state[0]->buf.as<uint8_t>(),
                           state[1]->buf.as<double>(), state[2]->buf.as<double>());
      break;
    case RG_VOP_INVERT:      // inverseNumPaths = 1/NumPaths, every vertex (BC-manyiter-v2.C:206-208)
      RG_TRY(rg::check_state(g, state, nstate, {8}, "vertexMap(invert)"));
      hipLaunchKernelGGL(rg::k_invert, rg::grid_for(ctx, g->n), 256, 0, ctx->stream, state[0]->buf.as<double>(),
                         g->n);
      break;
    case RG_VOP_BC_FINAL:    // Dep = (Dep - inv)/inv, every vertex (BC-manyiter-v2.C:230-232)
      RG_TRY(rg::check_state(g, state, nstate, {8, 8}, "vertexMap(BC final)"));
      hipLaunchKernelGGL(rg::k_bc_final, rg::grid_for(ctx, g->n), 256, 0, ctx->stream, state[0]->buf.as<double>(),
                         state[1]->buf.as<double>(), g->n);
      break;
    case RG_VOP_ABS_DIFF:    // x = |x - y| (PageRank.C:125-127)
    case RG_VOP_ABS:         // x = |y| (PageRankDelta.C:136-137)
      RG_TRY(rg::check_state(g, state, nstate, {8, 8}, "vertexMap(abs)"));
      hipLaunchKernelGGL(rg::k_abs_diff, grid, 256, 0, ctx->stream, ids, cnt, s->all, state[0]->buf.as<double>(),
                         state[1]->buf.as<const double>(), vop == RG_VOP_ABS_DIFF ? 1 : 0);
      break;
    default:
      return set_error(RG_EINVAL, "vertexMap: unknown operator");
  }
  RG_HIP(hipGetLastError());
  return RG_OK;
}

// vertexFilter (ligra.h:500-509) with the reference's filter functors.
int rg_vertexfilter(rg_graph* g, int fop, rg_subset* s, rg_array* const* state, int nstate, const double* params,
                    int nparams, rg_subset* out) {
  if (!g || !s || !out || s->g->n != g->n || out->g->n != g->n)
    return set_error(RG_EINVAL, "vertexFilter: bad subset");
  if (nparams < 0 || (nparams > 0 && !params)) return set_error(RG_EINVAL, "vertexFilter: bad parameters");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  const int64_t n = g->n;
  const int64_t cnt = s->all ? n : s->size;
  if (out == s && !s->all) return set_error(RG_EINVAL, "vertexFilter: output must differ from a listed input");
  // Keep flags: a byte per vertex, all zero between calls.
  rg::DevBuf& kb = ctx->scratch[10];
  if (kb.bytes < (size_t)std::max<int64_t>(n, 1) || !kb.p) {
    RG_TRY(kb.ensure((size_t)std::max<int64_t>(n, 1)));
    RG_TRY(rg::fill<uint8_t>(ctx, kb.as<uint8_t>(), (int64_t)kb.bytes, (uint8_t)0));
  }
  uint8_t* keep = kb.as<uint8_t>();
  const uint32_t* ids = s->ids.as<uint32_t>();
  const int grid = rg::grid_for(ctx, std::max<int64_t>(cnt, 1));
  auto need = [&](int k, const char* who) -> int {
    return nparams == k ? RG_OK : set_error(RG_EINVAL, std::string(who) + ": " + std::to_string(k) + " parameters");
  };
  if (cnt > 0) switch (fop) {
    case RG_FOP_PRD_FIRST:   // {p, Delta, nghSum}; params {damping, one_over_n, epsilon2}
      RG_TRY(rg::check_state(g, state, nstate, {8, 8, 8}, "vertexFilter(PR_Vertex_F_FirstRound)"));
      RG_TRY(need(3, "vertexFilter(PR_Vertex_F_FirstRound)"));
      hipLaunchKernelGGL(rg::k_filter_prd, grid, 256, 0, ctx->stream, ids, cnt, s->all, state[0]->buf.as<double>(),
                         state[1]->buf.as<double>(), state[2]->buf.as<const double>(), params[0],
                         (1.0 - params[0]) * params[1], params[1], params[2], 1, keep);
      break;
    case RG_FOP_PRD:         // {p, Delta, nghSum}; params {damping, epsilon2}
      RG_TRY(rg::check_state(g, state, nstate, {8, 8, 8}, "vertexFilter(PR_Vertex_F)"));
      RG_TRY(need(2, "vertexFilter(PR_Vertex_F)"));
      hipLaunchKernelGGL(rg::k_filter_prd, grid, 256, 0, ctx->stream, ids, cnt, s->all, state[0]->buf.as<double>(),
                         state[1]->buf.as<double>(), state[2]->buf.as<const double>(), params[0], 0.0, 0.0,
                         params[1], 0, keep);
      break;
    case RG_FOP_U32_EQ:      // {x u32}; params {value}
      RG_TRY(rg::check_state(g, state, nstate, {4}, "vertexFilter(x == value)"));
      RG_TRY(need(1, "vertexFilter(x == value)"));
      if (!(params[0] >= 0.0 && params[0] <= 4294967295.0 && params[0] == std::floor(params[0])))
        return set_error(RG_EINVAL, "vertexFilter(x == value): value must be a 32-bit unsigned integer");
      hipLaunchKernelGGL(rg::k_filter_u32_eq, grid, 256, 0, ctx->stream, ids, cnt, s->all,
                         state[0]->buf.as<const uint32_t>(), (uint32_t)params[0], keep);
      break;
    case RG_FOP_U8_NONZERO:  // {x u8}
      RG_TRY(rg::check_state(g, state, nstate, {1}, "vertexFilter(x != 0)"));
      RG_TRY(need(0, "vertexFilter(x != 0)"));
      hipLaunchKernelGGL(rg::k_filter_u8_nz, grid, 256, 0, ctx->stream, ids, cnt, s->all,
                         state[0]->buf.as<const uint8_t>(), keep);
      break;
    default:
      return set_error(RG_EINVAL, "vertexFilter: unknown operator");
  } else if (fop < RG_FOP_PRD_FIRST || fop > RG_FOP_U8_NONZERO) {
    return set_error(RG_EINVAL, "vertexFilter: unknown operator");
  }
  RG_HIP(hipGetLastError());
  // Compact the flags (all n) into an ascending list.
  const int64_t chunk = std::max<int64_t>(rg::kCompactBlock, (n + rg::kCompactBlocks - 1) / rg::kCompactBlocks);
  const int nb = (int)std::max<int64_t>(1, (n + chunk - 1) / chunk);
  int64_t* offs = nullptr;
  rg::DevBuf& ob = ctx->scratch[11];
  RG_TRY(ob.ensure(sizeof(int64_t) * (rg::kCompactBlocks + 1)));
  offs = ob.as<int64_t>();
  RG_TRY(out->ids.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(n, 1)));
  hipLaunchKernelGGL(rg::k_compact_count, nb, rg::kCompactBlock, 0, ctx->stream, keep, n, chunk, offs);
  hipLaunchKernelGGL(rg::k_compact_scan, 1, rg::kCompactBlocks, 0, ctx->stream, offs, nb);
  hipLaunchKernelGGL(rg::k_compact_write, nb, rg::kCompactBlock, 0, ctx->stream, keep, n, chunk, offs,
                     out->ids.as<uint32_t>());
  RG_HIP(hipGetLastError());
  int64_t total = 0;
  RG_TRY(rg::ctx_read_count(ctx, offs + nb, &total, 1));
  out->size = total;
  out->degsum = -1;
  out->all = false;
  return RG_OK;
}

// rg_bfs_noalloc's layout of the fixed scratch array: 32 int64 counters,
// two frontier lists, the byte map and the heavy queue, each 256-byte
// aligned.
namespace {
struct FixedLayout {
  size_t ids0, ids1, dense, heavy, end;
  int64_t heavy_cap;
};
FixedLayout fixed_layout(int64_t n, int64_t m, int heavy_deg) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t cnt = (size_t)std::max<int64_t>(n, 64);
  FixedLayout L;
  L.ids0 = 256;
  L.ids1 = L.ids0 + up(sizeof(uint32_t) * cnt);
  L.dense = L.ids1 + up(sizeof(uint32_t) * cnt);
  L.heavy = L.dense + up(cnt);
  L.heavy_cap = rg::heavy_queue_items(m, heavy_deg);
  L.end = L.heavy + up(sizeof(uint64_t) * (size_t)L.heavy_cap);
  return L;
}
}  // namespace

int rg_bfs_noalloc(rg_graph* g, uint32_t src, uint32_t* parents_dev, rg_bfs_summary* out) {
  if (!g || !parents_dev || !out) return set_error(RG_EINVAL, "rg_bfs_noalloc: null argument");
  rg_ctx* ctx = g->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  if ((int64_t)src >= g->n) return set_error(RG_EINVAL, "rg_bfs_noalloc: source out of range");
  if (ctx->timing) return set_error(RG_EINVAL, "rg_bfs_noalloc: per-launch timing is on (it creates events)");
  if (!ctx->fixed_scratch) return set_error(RG_EDEVICE, "rg_bfs_noalloc: no fixed scratch on this device");
  rg_opts o;
  rg_opts_default(&o);
  o.variant = RG_PUSH;
  const FixedLayout L = fixed_layout(g->n, g->m, 1024);
  if (L.end > (size_t)RG_FIXED_SCRATCH_BYTES)
    return set_error(RG_EINVAL, "rg_bfs_noalloc: this graph needs " + std::to_string(L.end) +
                                    " bytes of scratch (9 per vertex, 8 per 683 arcs), more than the fixed " +
                                    std::to_string((size_t)RG_FIXED_SCRATCH_BYTES));
  char* base = static_cast<char*>(ctx->fixed_scratch);
  rg::BfsWork w;
  w.counts = reinterpret_cast<int64_t*>(base);
  w.ids[0] = reinterpret_cast<uint32_t*>(base + L.ids0);
  w.ids[1] = reinterpret_cast<uint32_t*>(base + L.ids1);
  w.dense = reinterpret_cast<uint8_t*>(base + L.dense);
  w.heavy = reinterpret_cast<uint64_t*>(base + L.heavy);
  w.heavy_cap = L.heavy_cap;
  // The array keeps the last call's contents: counters and the byte map
  // start from zero (a call that stopped on an error may have left them set).
  RG_TRY(rg::fill<unsigned long long>(ctx, reinterpret_cast<unsigned long long*>(w.counts), 32, 0ull));
  RG_TRY(rg::fill<uint8_t>(ctx, w.dense, std::max<int64_t>(g->n, 64), (uint8_t)0));
  uint64_t cs = 0;
  RG_TRY(rg::bfs_push(g, src, &o, nullptr, &out->iters, &out->reached, &out->edges, nullptr, parents_dev, &cs,
                      &w));
  out->checksum = cs;
  return RG_OK;
}

// sequence::plusReduce over an FP64 state array, in a fixed order.
int rg_array_sum(const rg_array* a, double* out) {
  if (!a || !out || a->elem != 8) return set_error(RG_EINVAL, "rg_array_sum: an array of doubles");
  rg_ctx* ctx = a->ctx;
  RG_HIP(hipSetDevice(ctx->device));
  const int np = (int)std::min<int64_t>(std::max<int64_t>(1, (a->count + 255) / 256), (int64_t)ctx->num_cus * 4);
  rg::DevBuf& pb = ctx->scratch[11];
  RG_TRY(pb.ensure(sizeof(double) * (np + 1)));
  hipLaunchKernelGGL(rg::k_sum_partial, np, 256, 0, ctx->stream, a->buf.as<const double>(), a->count,
                     pb.as<double>());
  hipLaunchKernelGGL(rg::k_sum_final, 1, 256, 0, ctx->stream, pb.as<double>(), np);
  RG_HIP(hipGetLastError());
  int64_t bits = 0;
  RG_TRY(rg::ctx_read_count(ctx, reinterpret_cast<const int64_t*>(pb.as<double>() + np), &bits, 1));
  std::memcpy(out, &bits, sizeof(double));
  return RG_OK;
}

// The device address of a state array (for entry points taking device
// pointers).
int rg_array_device_ptr(const rg_array* a, void** out) {
  if (!a || !out) return set_error(RG_EINVAL, "rg_array_device_ptr: null");
  *out = a->buf.p;
  return RG_OK;
}

}  // extern "C"
