// Internal entry points of the query paths (single-source BFS, batched
// multi-source BFS, BC, PageRank) shared by the C ABI, the autotuner and the
// serving loop, and the per-level hash of the BFS level checksum.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace rg {

// splitmix64 finaliser (Steele, Lea, Flood 2014) of (level << 32 | v); a BFS
// query's checksum is the sum of level_hash(v, level(v)) over the reached
// vertices, modulo 2^64.
__host__ __device__ inline uint64_t level_hash(uint32_t v, uint32_t level) {
  uint64_t z = (((uint64_t)level << 32) | v) + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Weight of vertex v in a result summary: 53 hash bits as a double in [0, 1).
__host__ __device__ inline double summary_weight(uint32_t v) {
  return (double)(level_hash(v, 0xffffffffu) >> 11) * 0x1p-53;
}

// Device buffers of one single-source push BFS supplied by the caller
// instead of the context pools: two frontier lists and a byte map (n
// entries each; the map all zero on entry, and left so), the heavy queue
// (heavy_cap chunk items) and 16 int64 counters.  With a workspace the BFS
// runs the push engine (no hub copies, no pull or blocked levels), records
// no events and allocates nothing; parents_dev must then be given.
struct BfsWork {
  uint32_t* ids[2];
  uint8_t* dense;
  uint64_t* heavy;
  int64_t heavy_cap;
  int64_t* counts;
};

int bfs_push(rg_graph* g, uint32_t src, const rg_opts* o, uint32_t* parents_out, int64_t* iters_out,
             int64_t* reached_out, int64_t* edges_out, double* ms_out, uint32_t* parents_dev,
             uint64_t* checksum_out, const BfsWork* work = nullptr);

int bfs_batch(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, int32_t* levels_out,
              rg_bfs_summary* out, double* ms_out);

// Per-level view of a running bit-parallel batch (device pointers), handed
// to a level hook after a level's kernels are launched and before its
// frontier words are cleared.  Levels are launched ahead of the host's
// checks, so the hook's kernels read *mode on the device and do nothing
// once it is kMsDone.
constexpr int64_t kMsDone = 2;
struct MsLevelView {
  const int64_t* mode;
  const int64_t* level;                  // level of the current frontier
  const int64_t* ncount;                 // size of the next frontier list
  const uint32_t* list_nxt;              // vertices some source reaches first at level + 1
  const unsigned long long* front_cur;   // per vertex: sources for which it is at level
  const unsigned long long* front_nxt;   // per vertex: sources for which it is at level + 1
  const unsigned long long* active_cur;  // bit per vertex: front_cur != 0
};
struct MsLevelHook {
  virtual ~MsLevelHook() = default;
  virtual int level(const MsLevelView& v) = 0;
};

// A one-word batch (1..64 sources) with a hook called at every level.
int bfs_batch_hooked(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, MsLevelHook* hook,
                     rg_bfs_summary* out);

// Batched single-source BC (bcbatch.hip): up to 64 sources, each answered
// as rg_bc with that one source.
int bc_batch(rg_graph* g, const uint32_t* sources, int nsrc, const rg_opts* o, double* deps_out,
             rg_bc_summary* out, double* ms_out);
// Device bytes one context holds for BC batches on an (n, m) graph.
int64_t bc_batch_bytes(int64_t n, int64_t m);

// Weighted sum of the finite entries of x (weights in [0, 1) from a fixed
// hash of the vertex id) and the count of non-finite ones (serve.hip).
int fp_summary(rg_ctx* ctx, const double* x, int64_t n, double* wsum, int64_t* nonfinite);
int radii_batch(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o, int32_t* radii_out,
                int64_t* rounds_out, double* ms_out);

// Builds the blocked in-edge layout the batches use (once per graph).
int bfs_batch_prepare(rg_graph* g);
// Allocates g's context's batch state for batches of up to nsrc sources
// (so that a first batch does not pay for it).
int bfs_batch_reserve(rg_graph* g, int nsrc);
// Likewise for BC batches (their state and the one-word BFS batch state).
int bc_batch_reserve(rg_graph* g);
// Device bytes one context holds for batches of nsrc sources on an
// (n, m) graph.
int64_t bfs_batch_bytes(int64_t n, int64_t m, int nsrc);

// Constants of a PageRankDelta vertexFilter; first = 1 in round 1.
struct PrdStep {
  double damping, added, one_over_n, eps2;
  int first;
};

// One dense PageRankDelta round on the blocked layout (blocked.hip): the
// edge pass over contrib and the vertexFilter, fused.  Updates p, writes
// every vertex's next quotient into contrib_next, appends the active
// vertices to out_ids, adds their count to info[0] (zeroed by the caller)
// and stores their out-degree sum in info[1] and the round's L1 norm in *l1.
// work holds the split blocks' partials and the workgroups' sums.
int blocked_prd_round(rg_graph* g, const rg_opts* o, const double* contrib, double* contrib_next, double* p,
                      const PrdStep& st, uint32_t* out_ids, int64_t* info, double* l1, const int64_t* zero,
                      DevBuf& work);

// Builds the blocked PageRank layout for these options (once per graph).
int pagerank_prepare(rg_graph* g, const rg_opts* o);
// Builds the all-vertices heavy-chunk list of the push engine.
int static_heavy_list(rg_graph* g, int heavy_deg);

// One dense BC level (forward path counts or backward dependencies) on the
// blocked layout of g's current orientation (bcdense.hip).
// active_bits (nullable): a complete bitmap of the frontier ids[0, count)
// the caller keeps (else the bits are set from ids and cleared after);
// next_bits (forward, nullable): receives the words of the next frontier's
// bitmap for every block with an unvisited destination (the caller zeroes
// it first).
int bc_dense_level(rg_graph* g, bool fwd, const uint32_t* ids, int64_t count, const double* val, double* out,
                   const uint8_t* visited, uint32_t* out_ids, int64_t* out_info,
                   const unsigned long long* active_bits = nullptr, unsigned long long* next_bits = nullptr);
// Sets the bits of ids[0, count) in `bits` (atomic OR per vertex).
int set_frontier_bits(rg_ctx* ctx, const uint32_t* ids, int64_t count, unsigned long long* bits);

// One dense BFS level (BFS_F: unvisited vertices with a frontier in-neighbour
// take one as parent and join the next frontier) on the blocked layout.
int bfs_dense_level(rg_graph* g, const uint32_t* ids, int64_t count, uint32_t* parents, uint32_t* out_ids,
                    int64_t* out_info);

// First in-neighbour of every vertex of the current orientation (the pull
// BFS's first probe), built on first use and kept with the graph.
int in_first_neighbours(rg_graph* g, const uint32_t** out);

// One bottom-up BFS level (the push-pull tree's pull edgeMap with early
// exit) over g's in-edges (bcdense.hip).  cur_bits: the complete frontier
// bitmap the previous pull level returned in *next_bits, or null (the
// frontier is then taken from ids).  *next_bits (nullable) receives the
// next frontier's complete bitmap, valid until the next dense call on the
// context.
int bfs_pull_level(rg_graph* g, const uint32_t* ids, int64_t count, uint32_t* parents, int heavy_deg,
                   uint32_t* out_ids, int64_t* out_info, const unsigned long long* cur_bits = nullptr,
                   const unsigned long long** next_bits = nullptr);

// Pull levels whose kind is decided on the device (the single-source BFS
// level loop run as a captured chunk): every level's kernels are launched,
// and those of a pull level run only when *mode == want.  chain: where the
// level's frontier is as a complete bitmap -- 0: only as the list (its bits
// are set into the zero bitmap and cleared after), 1: full0, 2: full1 (the
// previous pull level's output); the level writes the other full bitmap.
struct PullSel {
  const int64_t* mode = nullptr;
  const int64_t* chain = nullptr;
  int64_t want = 0;
  const unsigned long long* zero_bits = nullptr;
  unsigned long long* full0 = nullptr;
  unsigned long long* full1 = nullptr;
};

// The pull levels' buffers (those bfs_pull_level uses), the bound of the
// vertices with in-edges (nn, wnz words) and the first in-neighbours.
struct PullBufs {
  unsigned long long* bits = nullptr;      // zero between levels
  unsigned long long* full[2] = {nullptr, nullptr};
  uint32_t* heavy = nullptr;
  int heavy_deg = 256;
  int64_t nn = 0, nwords = 0, wnz = 0;
  const uint32_t* first_in = nullptr;
};
int bfs_pull_buffers(rg_graph* g, int heavy_deg, PullBufs* pb);
// Zeroes the full bitmaps' words past the vertices with in-edges (no pull
// level writes them), once before a device-controlled loop.
int bfs_pull_zero_tails(rg_ctx* ctx, const PullBufs& pb);
// Launches one device-controlled pull level: frontier list ids with its
// size at *count_dev, heavy counter *hcount (zero when the level starts),
// next frontier appended to out_ids with out_info[0..1] += count, degree sum.
int bfs_pull_level_dev(rg_graph* g, const PullBufs& pb, const PullSel& sel, const uint32_t* ids,
                       const int64_t* count_dev, uint32_t* parents, int64_t* hcount, uint32_t* out_ids,
                       int64_t* out_info);

int bc_push(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
            int64_t* combined_out, double* ms_out);

}  // namespace rg
