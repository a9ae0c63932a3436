// Query serving over a resident graph: a request loop that admits BFS, BC
// and PageRank queries by arrival time, runs them on several device
// contexts ("slots", one stream each) that share the graph's arrays, and
// coalesces the BFS queries that have arrived into bit-parallel batches of
// up to 256 and the BC queries into batches of up to 64 (continuous
// batching: a query that arrives while batches are running joins the batch
// the next free slot starts).  The reference runs one
// Compute per process and has no such loop; the per-query semantics are its
// apps' (BFS.C, BC-manyiter-v2.C for one source, PageRank.C).
//
// Host side only: one std::thread per slot pulls work from shared pending
// queues under a mutex and drives the synchronous query entry points on
// its own context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "blocked.h"
#include "device.h"
#include "engine_host.h"
#include "queries.h"

namespace rg {

constexpr int kSumBlocks = 512;
constexpr int kBcBatch = 64;   // BC queries per batch (one word of source bits)
// Batch widths timed before the clock starts, per batched kind.  The cost
// of a bit-parallel batch is far from linear in its width: on a symmetric
// R-MAT of 2^26 vertices a BFS batch of 1 / 8 / 64 / 256 sources takes
// 17 / 32 / 37 / 120 ms against 4 ms per single-source pull query, a BC
// batch of 1 / 8 / 64 takes 99 / 273 / 489 ms against 22 ms per
// single-source query.
constexpr int kCalWidths[] = {1, 8, 64, 256};

// Weighted sum of the finite entries (weights w(v) in [0,1) from a fixed
// hash of v) and count of the non-finite ones: per-block partial sums in a
// fixed tree order, summed on the host in block order -- deterministic for
// a given array.
__global__ __launch_bounds__(256) void k_fp_summary(const double* x, int64_t n, double* part,
                                                    unsigned long long* nonfinite) {
  __shared__ double red[256];
  double s = 0.0;
  unsigned long long nf = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const double xv = x[v];
    if (!isfinite(xv)) {
      nf++;
    } else {
      s += xv * summary_weight((uint32_t)v);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
  for (int o = 32; o > 0; o >>= 1) nf += __shfl_xor(nf, o);
  if (__lane_id() == 0 && nf) atomicAdd(nonfinite, nf);
}

int fp_summary(rg_ctx* ctx, const double* x, int64_t n, double* wsum, int64_t* nonfinite) {
  if (!x) return set_error(RG_EDEVICE, "serve: no result array");
  DevBuf& b = ctx->scratch[5];
  RG_TRY(b.ensure(sizeof(double) * (kSumBlocks + 1)));
  double* part = b.as<double>();
  unsigned long long* nf = reinterpret_cast<unsigned long long*>(part + kSumBlocks);
  RG_TRY(fill<unsigned long long>(ctx, nf, 1, 0ull));
  hipLaunchKernelGGL(k_fp_summary, kSumBlocks, 256, 0, ctx->stream, x, n, part, nf);
  RG_HIP(hipGetLastError());
  std::vector<double> h(kSumBlocks + 1);
  RG_TRY(ctx_copy(ctx, h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (int i = 0; i < kSumBlocks; ++i) s += h[i];
  uint64_t c;
  std::memcpy(&c, &h[kSumBlocks], sizeof(c));
  *wsum = s;
  *nonfinite = (int64_t)c;
  return RG_OK;
}

namespace {

struct Slot {
  rg_ctx* ctx = nullptr;
  rg_graph* view = nullptr;
};

struct Shared {
  const rg_query* q;
  int64_t nq;
  std::vector<int64_t> order;      // queries by arrival time
  int64_t next = 0;                // next query (in order) not yet arrived
  std::deque<int64_t> pending[3];
  int64_t batches = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::chrono::steady_clock::time_point t0;
  int err = RG_OK;
  std::string msg;
  // Time of one single-source run and of batches of the calibrated widths,
  // per kind, measured alone before the clock starts.  k waiting queries of
  // a kind are batched when the batch ends no later than the last of them
  // would as single-source runs: cost(k) <= k * single.  The runs are bound
  // by device memory bandwidth, so slots running side by side share it and
  // do not change the comparison.
  double single_ms[3] = {0.0, 0.0, 0.0};
  std::vector<std::pair<int, double>> curve[3];   // (width, ms), widths ascending
  // Batch time at width k, linear between the timed widths and beyond the
  // last two.
  double batch_cost(int kind, size_t k) const {
    const auto& c = curve[kind];
    if (c.empty()) return 0.0;
    if (c.size() == 1) return c[0].second * (double)k / (double)c[0].first;
    size_t i = 1;
    while (i + 1 < c.size() && (size_t)c[i].first < k) ++i;
    const double x0 = c[i - 1].first, y0 = c[i - 1].second, x1 = c[i].first, y1 = c[i].second;
    return y0 + (y1 - y0) * ((double)k - x0) / (x1 - x0);
  }
  bool batch_pays(int kind, size_t k) const {
    return k >= 2 && batch_cost(kind, k) <= (double)k * single_ms[kind];
  }
  double ms_now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  // Admission by device memory.  Each slot's context holds device memory
  // for query state: its batched BFS state (ms), its batched BC state (bcb)
  // and its pools (base), measured after every item it runs.  An item
  // starts on a slot only when everything the slots hold, plus the growth
  // the items in flight may still need, plus this item's growth beyond what
  // its slot holds that it can reuse, stays within the budget; otherwise a
  // narrower batch is tried, batch state the item does not use is dropped
  // (on this slot, then on idle ones), or the item waits until an item in
  // flight ends.
  int64_t budget = 0;
  int64_t n = 0, m = 0, pr_bytes = 0;
  Slot* slots = nullptr;
  int nslots = 0;
  std::vector<int64_t> held_ms, held_bcb, held_base, growth;
  std::vector<char> busy, drop;   // drop: an idle slot is asked to free its batch state
  std::vector<int> running;       // kind of the item a busy slot runs
  int64_t peak = 0, mem_waits = 0;
  // Items wait for memory: frees (which synchronise the device) happen only
  // then, never on a path every item takes.
  int mem_short = 0;
  int64_t held_total() const {
    int64_t t = 0;
    for (int i = 0; i < nslots; ++i) t += held_ms[i] + held_bcb[i] + held_base[i] + (busy[i] ? growth[i] : 0);
    return t;
  }
  void measure(int i) {
    const rg_ctx* ctx = slots[i].ctx;
    held_ms[i] = ms_state_bytes(ctx->ms);
    held_bcb[i] = bcb_state_bytes(ctx->bcb);
    held_base[i] = ctx_device_bytes(ctx) - held_ms[i] - held_bcb[i];
  }
  // Device bytes of query state one work item needs on its context.
  int64_t item_bytes(int kind, size_t k) const {
    const int64_t heavy = (m / 2048 + m / 1024 + 64) * 8;
    if (kind == RG_Q_BFS) return k > 1 ? bfs_batch_bytes(n, m, (int)k) + heavy : 14 * n + heavy;
    if (kind == RG_Q_BC) return k > 1 ? bfs_batch_bytes(n, m, kBcBatch) + bc_batch_bytes(n, m) + heavy : n * 38 + heavy;
    return pr_bytes;
  }
  // What the item may add to slot i: its bytes beyond what the slot holds
  // and the item can use (batch items use the batch states).
  int64_t item_growth(int i, int kind, size_t k) const {
    int64_t reuse = held_base[i];
    if (k > 1) reuse += held_ms[i] + (kind == RG_Q_BC ? held_bcb[i] : 0);
    return std::max<int64_t>(0, item_bytes(kind, k) - reuse);
  }
};

// Picks the next work item: the kind whose oldest pending query arrived
// first; the waiting BFS or BC queries as one batch when that pays.
// taken_ms: the item's start on the trace's clock, read under the lock (so
// items start in the order they are taken).  Returns false when every
// query has been taken.
bool take_work(Shared& S, const rg_serve_config& c, int slot, int* kind, std::vector<int64_t>& work,
               int64_t* batch, double* taken_ms) {
  std::unique_lock<std::mutex> lk(S.mu);
  bool waited = false;
  for (;;) {
    if (S.err != RG_OK) return false;
    if (S.drop[slot]) {   // asked while idle: free this slot's batch state
      ctx_release_batch_state(S.slots[slot].ctx, 3);
      S.measure(slot);
      S.drop[slot] = 0;
    }
    const double now = S.ms_now();
    while (S.next < S.nq && S.q[S.order[S.next]].arrival_ms <= now) {
      const int64_t i = S.order[S.next++];
      S.pending[S.q[i].kind].push_back(i);
    }
    int best = -1;
    double ba = 0.0;
    for (int k = 0; k < 3; ++k)
      if (!S.pending[k].empty()) {
        const double a = S.q[S.pending[k].front()].arrival_ms;
        if (best < 0 || a < ba) {
          best = k;
          ba = a;
        }
      }
    // Queries that run alone: none starts while one runs, and one starts
    // only when no other slot is busy.
    bool alone_running = false, any_busy_slot = false;
    for (int j = 0; j < S.nslots; ++j)
      if (S.busy[j]) {
        any_busy_slot = true;
        alone_running = alone_running || ((c.alone_kinds >> S.running[j]) & 1);
      }
    if (best >= 0 && (alone_running || (((c.alone_kinds >> best) & 1) && any_busy_slot))) {
      S.cv.wait_for(lk, std::chrono::milliseconds(2));
      continue;
    }
    if (best >= 0) {
      size_t take = 1;
      // BFS and BC: the waiting queries of the kind that arrived before the
      // oldest waiting query of any other kind (up to the batch limit) as one
      // batch when that answers them sooner than single-source runs would,
      // else the oldest alone.  A batch never takes queries that arrived
      // after an older query of another kind, so batching keeps the order of
      // arrival across kinds (later BFS queries riding in a batch would
      // otherwise overtake an older BC query by the batch's length).
      if ((best == RG_Q_BFS || best == RG_Q_BC) && !c.one_at_a_time) {
        const size_t cap = best == RG_Q_BFS ? (size_t)c.max_batch : (size_t)std::min(c.max_batch, kBcBatch);
        double other = std::numeric_limits<double>::infinity();
        for (int k2 = 0; k2 < 3; ++k2)
          if (k2 != best && !S.pending[k2].empty()) other = std::min(other, S.q[S.pending[k2].front()].arrival_ms);
        size_t k = 0;
        for (int64_t qi : S.pending[best]) {
          if (k >= cap || S.q[qi].arrival_ms > other) break;
          ++k;
        }
        if (S.batch_pays(best, k)) take = k;
      }
      // Admission by memory: the widest batch that pays and fits, else
      // drop unused batch state, else wait for an item in flight to end.
      const int64_t others = S.held_total() - S.held_ms[slot] - S.held_bcb[slot] - S.held_base[slot];
      auto fits = [&](size_t k) {
        return others + S.held_ms[slot] + S.held_bcb[slot] + S.held_base[slot] + S.item_growth(slot, best, k) <=
               S.budget;
      };
      bool narrowed = false;
      while (take > 1 && !fits(take)) {
        const size_t half = take / 2;
        take = half >= 2 && S.batch_pays(best, half) ? half : 1;
        narrowed = true;
      }
      if (!fits(take)) {
        // Batch state this item does not use, on this slot first.
        const int unused = take == 1 ? 3 : (best == RG_Q_BFS ? 2 : 0);
        const bool mine = (unused & 1 && S.held_ms[slot]) || (unused & 2 && S.held_bcb[slot]);
        if (mine) {
          ctx_release_batch_state(S.slots[slot].ctx, unused);
          S.measure(slot);
          continue;
        }
        bool asked = false, any_busy = false;
        for (int j = 0; j < S.nslots; ++j) {
          if (j == slot) continue;
          any_busy = any_busy || S.busy[j];
          if (!S.busy[j] && (S.held_ms[j] || S.held_bcb[j]) && !S.drop[j]) {
            S.drop[j] = 1;
            asked = true;
          }
        }
        if (asked || any_busy) {
          if (!waited) {
            S.mem_waits++;
            S.mem_short++;
          }
          waited = true;
          S.cv.notify_all();
          S.cv.wait_for(lk, std::chrono::milliseconds(5));
          continue;
        }
        // Nothing else holds or will free memory: the item runs alone.
      }
      if (waited) S.mem_short--;
      if (narrowed && !waited) S.mem_waits++;
      S.growth[slot] = S.item_growth(slot, best, take);
      S.busy[slot] = 1;
      S.running[slot] = best;
      S.peak = std::max(S.peak, S.held_total());
      for (size_t i = 0; i < take; ++i) {
        work.push_back(S.pending[best].front());
        S.pending[best].pop_front();
      }
      *kind = best;
      *batch = S.batches++;
      *taken_ms = S.ms_now();   // the item's start, under the lock
      bool more = false;
      for (auto& p : S.pending) more = more || !p.empty();
      if (more) S.cv.notify_all();
      return true;
    }
    if (S.next >= S.nq) {
      // Every query is taken: this slot's batch state is freed if items on
      // the other slots wait for memory.
      if (S.mem_short) {
        ctx_release_batch_state(S.slots[slot].ctx, 3);
        S.measure(slot);
        S.cv.notify_all();
      }
      return false;
    }
    S.cv.wait_until(lk, S.t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                   std::chrono::duration<double, std::milli>(S.q[S.order[S.next]].arrival_ms)));
  }
}

// Batches of several slots overlap on the device only as per-level kernels
// (one persistent batch runs at a time), so with more than one slot the
// library's choice of launch form becomes per-level kernels.
void serve_batch_launch(const rg_serve_config& c, rg_opts& qo) {
  if (qo.batch_launch == 0 && c.slots > 1) qo.batch_launch = 1;
}

int run_work(Slot& sl, const rg_serve_config& c, const rg_query* q, int kind, const std::vector<int64_t>& work,
             rg_query_result* res) {
  rg_graph* g = sl.view;
  rg_opts o = c.opts;
  rg_opts qo = c.opts;   // BFS / BC: the push engine, RADAR hub copies or blocked dense BC levels
  if (qo.variant == RG_PULL) qo.variant = RG_PUSH;
  serve_batch_launch(c, qo);
  const size_t k = work.size();
  std::vector<rg_bfs_summary> sum(k, rg_bfs_summary{0, 0, 0, 0});
  std::vector<int64_t> iters(k, 0);
  std::vector<double> l1(k, 0.0), wsum(k, 0.0);
  std::vector<int64_t> nonfin(k, 0);
  double ms = 0.0;
  if (kind == RG_Q_BFS) {
    if (k == 1) {
      uint64_t cs = 0;
      RG_TRY(bfs_push(g, q[work[0]].source, &c.bfs_opts, nullptr, &sum[0].iters, &sum[0].reached, &sum[0].edges,
                      &ms, nullptr, &cs));
      sum[0].checksum = cs;
    } else {
      std::vector<uint32_t> src(k);
      for (size_t i = 0; i < k; ++i) src[i] = q[work[i]].source;
      RG_TRY(bfs_batch(g, src.data(), (int)k, &qo, nullptr, sum.data(), &ms));
    }
  } else if (kind == RG_Q_BC && k > 1) {
    std::vector<uint32_t> src(k);
    for (size_t i = 0; i < k; ++i) src[i] = q[work[i]].source;
    std::vector<rg_bc_summary> bs(k);
    RG_TRY(bc_batch(g, src.data(), (int)k, &qo, nullptr, bs.data(), &ms));
    for (size_t i = 0; i < k; ++i) {
      iters[i] = bs[i].rounds;
      wsum[i] = bs[i].wsum;
      nonfin[i] = bs[i].nonfinite;
    }
  } else if (kind == RG_Q_BC) {
    const bool tr = g->transposed;
    const uint32_t s = q[work[0]].source;
    RG_TRY(bc_push(g, &s, 1, &qo, nullptr, &iters[0], &ms));
    if (g->transposed != tr) RG_TRY(rg_graph_transpose(g));
    RG_TRY(fp_summary(sl.ctx, sl.ctx->last_result, g->n, &wsum[0], &nonfin[0]));
  } else {
    const int64_t it = q[work[0]].iters > 0 ? q[work[0]].iters : 10;
    RG_TRY(rg_pagerank(g, it, &o, nullptr, &iters[0], &l1[0], &ms));
    RG_TRY(fp_summary(sl.ctx, sl.ctx->last_result, g->n, &wsum[0], &nonfin[0]));
  }
  for (size_t i = 0; i < k; ++i) {
    rg_query_result& R = res[work[i]];
    R.bfs = sum[i];
    R.iters = iters[i];
    R.l1 = l1[i];
    R.wsum = wsum[i];
    R.nonfinite = nonfin[i];
  }
  return RG_OK;
}

// Before the clock starts: the time of one single-source run and of
// batches of the calibrated widths (up to the batch limit) of each batched
// kind in the trace, on the kind's own sources (repeated when the trace
// holds fewer); answers are discarded.  The single-source run and the
// first batch go once untimed first (scratch set-up).
int calibrate(Shared& S, Slot& sl, const rg_serve_config& c, const rg_query* q, int64_t nq) {
  using clock = std::chrono::steady_clock;
  rg_graph* g = sl.view;
  rg_opts qo = c.opts;
  if (qo.variant == RG_PULL) qo.variant = RG_PUSH;
  serve_batch_launch(c, qo);
  for (int kind : {(int)RG_Q_BFS, (int)RG_Q_BC}) {
    std::vector<uint32_t> srcs;
    for (int64_t i = 0; i < nq; ++i)
      if (q[i].kind == kind) srcs.push_back(q[i].source);
    if (srcs.empty()) continue;
    const int cap = kind == RG_Q_BFS ? c.max_batch : std::min(c.max_batch, kBcBatch);
    auto single = [&]() -> int {
      double ms = 0.0;
      if (kind == RG_Q_BFS) {
        int64_t it = 0, reached = 0, edges = 0;
        return bfs_push(g, srcs[0], &c.bfs_opts, nullptr, &it, &reached, &edges, &ms, nullptr, nullptr);
      }
      const bool tr = g->transposed;
      int64_t it = 0;
      RG_TRY(bc_push(g, &srcs[0], 1, &qo, nullptr, &it, &ms));
      return g->transposed != tr ? rg_graph_transpose(g) : RG_OK;
    };
    RG_TRY(single());
    auto t = clock::now();
    RG_TRY(single());
    S.single_ms[kind] = std::chrono::duration<double, std::milli>(clock::now() - t).count();
    for (int w : kCalWidths) {
      const int k = std::min(w, cap);
      if (!S.curve[kind].empty() && S.curve[kind].back().first >= k) break;
      std::vector<uint32_t> b(k);
      for (int i = 0; i < k; ++i) b[i] = srcs[i % srcs.size()];
      std::vector<rg_bfs_summary> bs(k);
      std::vector<rg_bc_summary> cs(k);
      double ms = 0.0;
      const int reps = S.curve[kind].empty() ? 2 : 1;   // batch state is reserved for every width
      for (int rep = 0; rep < reps; ++rep) {
        t = clock::now();
        if (kind == RG_Q_BFS) RG_TRY(bfs_batch(g, b.data(), k, &qo, nullptr, bs.data(), &ms));
        else RG_TRY(bc_batch(g, b.data(), k, &qo, nullptr, cs.data(), &ms));
      }
      S.curve[kind].emplace_back(k, std::chrono::duration<double, std::milli>(clock::now() - t).count());
    }
  }
  return RG_OK;
}

int warm_slot(Slot& sl, const rg_serve_config& c, const rg_query* q, int64_t nq) {
  bool done[3] = {false, false, false};
  std::vector<rg_query_result> scratch_res(1);
  for (int64_t i = 0; i < nq; ++i) {
    const int kind = q[i].kind;
    if (done[kind]) continue;
    done[kind] = true;
    // A private copy of the query so the trace's results are untouched.
    rg_query one = q[i];
    const std::vector<int64_t> work = {0};
    RG_TRY(run_work(sl, c, &one, kind, work, scratch_res.data()));
  }
  return RG_OK;
}

}  // namespace

}  // namespace rg

using rg::set_error;

extern "C" {

void rg_serve_config_default(rg_serve_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->slots = 2;
  c->max_batch = 256;
  c->one_at_a_time = 0;
  c->hbm_reserve = (int64_t)4 << 30;
  rg_opts_default(&c->opts);
  c->opts.variant = RG_BLOCKED;
  rg_opts_default(&c->bfs_opts);
  c->bfs_opts.variant = RG_PULL;
  c->alone_kinds = 1 << RG_Q_PAGERANK;
}

int rg_serve(rg_graph* g, const rg_serve_config* cfg, const rg_query* queries, int64_t nq,
             rg_query_result* results, rg_serve_stats* stats) {
  if (!g || !queries || !results || nq < 0) return set_error(RG_EINVAL, "rg_serve: bad arguments");
  rg_serve_config c;
  if (cfg) c = *cfg;
  else rg_serve_config_default(&c);
  if (c.slots <= 0) c.slots = 2;
  if (c.max_batch <= 0 || c.max_batch > 256) c.max_batch = 256;
  if (c.hbm_reserve <= 0) c.hbm_reserve = (int64_t)4 << 30;
  if (c.one_at_a_time) {
    c.slots = 1;
    c.max_batch = 1;
  }
  for (int64_t i = 0; i < nq; ++i) {
    const rg_query& q = queries[i];
    if (q.kind < RG_Q_BFS || q.kind > RG_Q_PAGERANK) return set_error(RG_EINVAL, "rg_serve: bad query kind");
    if (q.kind != RG_Q_PAGERANK && (int64_t)q.source >= g->n)
      return set_error(RG_EINVAL, "rg_serve: query source out of range");
    if (!(q.arrival_ms >= 0.0)) return set_error(RG_EINVAL, "rg_serve: negative arrival time");
  }
  const int dev = g->ctx->device;
  RG_HIP(hipSetDevice(dev));
  // Layouts every slot reads, built once on the shared graph.
  bool any_pr = false, any_bfs = false, any_bc = false;
  for (int64_t i = 0; i < nq; ++i) {
    any_pr |= queries[i].kind == RG_Q_PAGERANK;
    any_bfs |= queries[i].kind == RG_Q_BFS;
    any_bc |= queries[i].kind == RG_Q_BC;
  }
  const bool bc_blocked = c.opts.variant == RG_RADAR || c.opts.variant == RG_BLOCKED;
  if (((any_bfs || any_bc) && !c.one_at_a_time) || (any_bc && bc_blocked)) RG_TRY(rg::bfs_batch_prepare(g));
  if (any_bfs && c.bfs_opts.variant == RG_PULL) {
    const uint32_t* first = nullptr;   // shared by the slots' views
    RG_TRY(rg::in_first_neighbours(g, &first));
  }
  if (any_pr && c.opts.variant == RG_BLOCKED) RG_TRY(rg::pagerank_prepare(g, &c.opts));
  if (any_pr && c.opts.variant != RG_BLOCKED && c.opts.variant != RG_PULL)
    RG_TRY(rg::static_heavy_list(g, c.opts.heavy_degree > 0 ? c.opts.heavy_degree : 1024));
  RG_HIP(hipDeviceSynchronize());
  // Admission by device memory: the slots' query state may hold free HBM
  // less the reserve; one slot must hold a full-width BFS batch plus the
  // larger of a BC and a PageRank query, or the configuration is refused.
  // Each item is then admitted against what the slots hold (take_work).
  size_t free_b = 0, total_b = 0;
  RG_HIP(hipMemGetInfo(&free_b, &total_b));
  const int64_t n = g->n, m = g->m;
  const int64_t bfs_b = rg::bfs_batch_bytes(n, m, c.max_batch);
  const int64_t bc_b = std::max<int64_t>(n * 38 + (m / 2048 + m / 1024 + 64) * 8,
                                         c.one_at_a_time ? 0 : rg::bc_batch_bytes(n, m));
  const int64_t pr_b = n * 24 + ((n >> 14) + m / rg::kUnitEdges + 2) * (int64_t)(16384 * 8);
  const int64_t slot_bytes = bfs_b + std::max(bc_b, pr_b);
  const int64_t budget = (int64_t)free_b - c.hbm_reserve;
  if (budget < slot_bytes)
    return set_error(RG_ENOMEM, "rg_serve: free device memory holds no slot (" + std::to_string(budget) +
                                    " bytes free beyond the reserve, " + std::to_string(slot_bytes) + " per slot)");
  const int nslots = std::max(1, c.slots);

  std::vector<rg::Slot> slots(nslots);
  int rc = RG_OK;
  int64_t reserved = 0;
  std::vector<char> warm(nslots, 0);   // slots whose full state fits: reserved and warmed before the clock
  for (int i = 0; i < nslots && rc == RG_OK; ++i) {
    rc = rg_ctx_create(dev, &slots[i].ctx);
    if (rc == RG_OK) {
      slots[i].view = new rg_graph(*g);
      slots[i].view->ctx = slots[i].ctx;
      slots[i].view->view = true;
      slots[i].view->base = g;
      // Batch state allocated before the clock starts, on the slots whose
      // full state fits the budget (the others allocate when admitted).
      if (reserved + slot_bytes <= budget) {
        if (!c.one_at_a_time && any_bfs) rc = rg::bfs_batch_reserve(slots[i].view, c.max_batch);
        if (rc == RG_OK && !c.one_at_a_time && any_bc) rc = rg::bc_batch_reserve(slots[i].view);
        reserved += slot_bytes;
        warm[i] = 1;
      }
    }
  }
  if (rc == RG_OK) rc = hipDeviceSynchronize() == hipSuccess ? RG_OK : set_error(RG_EDEVICE, "rg_serve: set-up");
  rg::Shared S;
  S.q = queries;
  S.nq = nq;
  S.budget = budget;
  S.n = n;
  S.m = m;
  S.pr_bytes = pr_b;
  S.slots = slots.data();
  S.nslots = nslots;
  for (auto* v : {&S.held_ms, &S.held_bcb, &S.held_base, &S.growth}) v->assign(nslots, 0);
  S.busy.assign(nslots, 0);
  S.drop.assign(nslots, 0);
  S.running.assign(nslots, 0);
  S.order.resize(nq);
  std::iota(S.order.begin(), S.order.end(), 0);
  std::stable_sort(S.order.begin(), S.order.end(),
                   [&](int64_t a, int64_t b) { return queries[a].arrival_ms < queries[b].arrival_ms; });
  if (rc == RG_OK && !c.one_at_a_time) rc = rg::calibrate(S, slots[0], c, queries, nq);
  // Every other slot whose full state fits the budget runs one query of
  // each kind in the trace before the clock starts, so that it allocates no
  // pools (which synchronises the device) while the trace runs; the others
  // allocate when admitted.
  for (int i = 1; i < nslots && rc == RG_OK; ++i)
    if (warm[i]) rc = rg::warm_slot(slots[i], c, queries, nq);
  if (rc == RG_OK) {
    if (hipDeviceSynchronize() != hipSuccess) rc = set_error(RG_EDEVICE, "rg_serve: calibration");
    for (int i = 0; i < nslots; ++i) S.measure(i);
    S.peak = S.held_total();
  }
  if (rc == RG_OK) {
    S.t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int i = 0; i < nslots; ++i) {
      th.emplace_back([&, i] {
        if (hipSetDevice(dev) != hipSuccess) {
          std::lock_guard<std::mutex> lk(S.mu);
          S.err = RG_EDEVICE;
          S.msg = "rg_serve: hipSetDevice failed in a slot thread";
          S.cv.notify_all();
          return;
        }
        for (;;) {
          int kind = 0;
          int64_t batch = 0;
          std::vector<int64_t> work;
          double start = 0.0;
          if (!rg::take_work(S, c, i, &kind, work, &batch, &start)) return;
          int r = rg::run_work(slots[i], c, queries, kind, work, results);
          const double end = S.ms_now();
          {
            std::lock_guard<std::mutex> lk(S.mu);   // the item's memory is now measured, not estimated
            S.busy[i] = 0;
            S.growth[i] = 0;
            S.measure(i);
            S.peak = std::max(S.peak, S.held_total());
            S.cv.notify_all();
          }
          if (r != RG_OK) {
            std::lock_guard<std::mutex> lk(S.mu);
            if (S.err == RG_OK) {
              S.err = r;
              S.msg = rg_last_error();
            }
            S.cv.notify_all();
            return;
          }
          for (int64_t qi : work) {
            rg_query_result& R = results[qi];
            R.kind = kind;
            R.slot = i;
            R.batch_id = batch;
            R.batch_size = (int64_t)work.size();
            R.arrival_ms = queries[qi].arrival_ms;
            R.start_ms = start;
            R.end_ms = end;
            R.latency_ms = end - queries[qi].arrival_ms;
          }
        }
      });
    }
    for (auto& t : th) t.join();
  }
  for (auto& sl : slots) {
    if (sl.view) rg_graph_free(sl.view);
    if (sl.ctx) rg_ctx_destroy(sl.ctx);
  }
  if (rc != RG_OK) return rc;
  if (S.err != RG_OK) return set_error(S.err, S.msg);
  if (stats) {
    std::vector<double> lat(nq);
    double mk = 0.0, tot = 0.0;
    int64_t nbfs = 0, nbc = 0;
    for (int64_t i = 0; i < nq; ++i) {
      lat[i] = results[i].latency_ms;
      mk = std::max(mk, results[i].end_ms);
      tot += lat[i];
      nbfs += results[i].kind == RG_Q_BFS;
      nbc += results[i].kind == RG_Q_BC;
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      int64_t r = (int64_t)std::ceil(p * (double)lat.size());
      return lat[std::min<int64_t>(std::max<int64_t>(r, 1), (int64_t)lat.size()) - 1];
    };
    std::memset(stats, 0, sizeof(*stats));
    stats->queries = nq;
    stats->batches = S.batches;
    stats->slots = nslots;
    stats->makespan_ms = mk;
    stats->qps = mk > 0 ? (double)nq / (mk * 1e-3) : 0.0;
    stats->p50_ms = pct(0.50);
    stats->p99_ms = pct(0.99);
    stats->mean_ms = nq ? tot / (double)nq : 0.0;
    int64_t bfs_batches = 0, bc_batches = 0;
    {
      std::vector<char> seen((size_t)S.batches, 0);
      for (int64_t i = 0; i < nq; ++i) {
        const int k = results[i].kind;
        if ((k == RG_Q_BFS || k == RG_Q_BC) && !seen[results[i].batch_id]) {
          seen[results[i].batch_id] = 1;
          (k == RG_Q_BFS ? bfs_batches : bc_batches)++;
        }
      }
    }
    stats->mean_batch = bfs_batches ? (double)nbfs / (double)bfs_batches : 0.0;
    stats->mean_bc_batch = bc_batches ? (double)nbc / (double)bc_batches : 0.0;
    stats->hbm_budget = budget;
    stats->slot_bytes = slot_bytes;
    stats->peak_bytes = S.peak;
    stats->mem_waits = S.mem_waits;
  }
  return RG_OK;
}

}  // extern "C"
