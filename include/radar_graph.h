/*
 * radar_graph.h -- C ABI of the MI355X-native RADAR-Graph traversal engine.
 *
 * The reference (CMUAbstract/RADAR-Graph) is five OpenMP copies of Ligra plus
 * GAP's triangle counting.  Its operator interface is Ligra's C++ template API
 * (edgeMap / vertexMap / vertexSubset with app functors) and each app's
 * Compute(graph&, commandLine, pvector<uintE>& new_ids).  Arbitrary C++
 * functors cannot cross a C ABI, so every functor of the reference maps to a
 * named device operator (rg_op) and every Compute body maps to one entry point
 * below.  All signatures use plain pointers and sizes; host pointers unless a
 * name says "dev".  Every function returns RG_OK (0) or an error code, with a
 * message in rg_last_error(); nothing here aborts the process.
 *
 * Reference citations are relative to the reference tree's src/ directory;
 * PNO = push-new-orig, PP = push-pull, DS = degree-sorting, HD = hubdup,
 * RD = RADAR.
 */
#ifndef RADAR_GRAPH_H
#define RADAR_GRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------- */
enum {
  RG_OK = 0,
  RG_EINVAL = 1,      /* bad argument (reference: abort/assert paths)          */
  RG_ENOMEM = 2,      /* host or device allocation failed                      */
  RG_EDEVICE = 3,     /* HIP runtime error                                      */
  RG_EIO = 4,         /* file could not be opened / read / written              */
  RG_EFORMAT = 5,     /* malformed input ("Bad input file", PNO IO.h:235-249)   */
  RG_EUNSUPPORTED = 6 /* operation not defined for this graph (e.g. directed TC)*/
};

/* Thread-local text of the last error on the calling thread. */
const char* rg_last_error(void);
/* Library version string. */
const char* rg_version(void);

/* ---- host graphs ---------------------------------------------------------
 * A CSR in host memory: uint64 offsets[n+1], uint32 neighbours[m]; for
 * directed graphs also the in-edge CSR (CSC).  Neighbour lists are sorted,
 * deduplicated and free of self-loops, which is what Ligra's text reader does
 * (PNO Ligra/ligra/IO.h:364-421 directed, 432-467 symmetric) and what GAP's
 * SquishCSR does (PNO GAP/src/builder.h:134-163).
 */
typedef struct rg_hgraph rg_hgraph;

enum { RG_FMT_ADJ = 0,        /* Ligra AdjacencyGraph text (IO.h:207-475)      */
       RG_FMT_ADJ_BINARY = 1, /* Ligra .config/.adj/.idx binary (IO.h:477-626) */
       RG_FMT_EL = 2,         /* GAP edge list text (reader.h:50-57)           */
       RG_FMT_SG = 3,         /* GAP serialized graph (reader.h:251-304)       */
       RG_FMT_WEL = 4,        /* GAP weighted edge list, weights dropped (reader.h:59-67) */
       RG_FMT_GR = 5,         /* DIMACS "a u v w" arcs (reader.h:69-83)          */
       RG_FMT_METIS = 6,      /* METIS .graph, 1-based (reader.h:85-141)         */
       RG_FMT_MTX = 7 };      /* Matrix Market coordinate (reader.h:143-217)     */

/* GAP generator (GAP/src/generator.h:64-124): 2^scale vertices,
 * degree*2^scale generated edges, R-MAT (A,B,C = 0.57,0.19,0.19) unless
 * `uniform`; seed kRandSeed = 27491095 (util.h:22).  Raw edge list, before
 * deduplication.  rg_generate_edge_count gives the array length. */
int64_t rg_generate_edge_count(int scale, int degree);
int rg_generate_edges(int scale, int degree, int uniform, uint32_t* src, uint32_t* dst);

/* Generated graph, built with the reader's dedup semantics.  symmetric=1
 * adds both directions (GAP -g / -u always symmetrize, command_line.h:77-78). */
int rg_hgraph_generate(int scale, int degree, int uniform, int symmetric, rg_hgraph** out);
/* Graph from an edge list (duplicates and self-loops removed). */
int rg_hgraph_from_edges(int64_t n, int64_t nedges, const uint32_t* src,
                         const uint32_t* dst, int symmetric, rg_hgraph** out);
/* Graph from an existing CSR.  dedup=1 applies the reader's sort/unique/
 * self-loop removal; dedup=0 takes the lists as they are (binary reader). */
int rg_hgraph_from_csr(int64_t n, int64_t m, const uint64_t* offsets,
                       const uint32_t* nbrs, int symmetric, int dedup, rg_hgraph** out);
int rg_hgraph_read(const char* path, int format, int symmetric, rg_hgraph** out);
/* Format by file suffix, as GAP's Reader dispatches (reader.h:219-249):
 * .el .wel .sg .gr .graph .mtx; anything else is Ligra text (RG_FMT_ADJ). */
int rg_format_of(const char* path);
int rg_hgraph_write(const rg_hgraph* g, const char* path, int format);
int rg_hgraph_info(const rg_hgraph* g, int64_t* n, int64_t* m, int* symmetric);
/* Borrowed views (valid until rg_hgraph_free).  in_edges=1 selects the CSC
 * of a directed graph (for symmetric graphs both views are the CSR). */
const uint64_t* rg_hgraph_offsets(const rg_hgraph* g, int in_edges);
const uint32_t* rg_hgraph_neighbors(const rg_hgraph* g, int in_edges);
void rg_hgraph_free(rg_hgraph* g);

/* ---- source pickers (host) ----------------------------------------------- */
/* GAP SourcePicker (benchmark.h:47-68): mt19937(27491095) +
 * uniform_int_distribution<int32>(0,n-1), rejecting out-degree 0. */
int rg_pick_sources_gap(const rg_hgraph* g, int count, uint32_t* out);
/* BC's picker (BC-manyiter-v2.C:149-168): mt19937(27491095) +
 * uniform_int_distribution<long>(0,n-1), rejecting out-degree 0 in the
 * orientation BC sees for that source (directed graphs alternate, because BC
 * transposes once per source).  Original ids. */
int rg_pick_sources_bc(const rg_hgraph* g, int count, uint32_t* out);
/* Ligra hashInt(ulong) (PNO Ligra/ligra/utils.h:347-355), used by Radii
 * (Radii.C:120-127) as hashInt(i) % n. */
uint64_t rg_hash_int(uint64_t a);

/* ---- device context ------------------------------------------------------ */
typedef struct rg_ctx rg_ctx;
int rg_device_count(int* count);
/* One context = one device + one HIP stream + preallocated work pools. */
int rg_ctx_create(int device, rg_ctx** out);
/* Free every rg_graph bound to a context before destroying the context. */
int rg_ctx_destroy(rg_ctx* ctx);
int rg_ctx_sync(rg_ctx* ctx);
/* Instrumentation (the reference times whole Compute calls with its Timer,
 * PNO Ligra/ligra/timer.h; this adds per-launch device time): with timing
 * enabled, every edge-map kernel launch is bracketed by HIP events on the
 * context stream.  rg_ctx_timing(ctx, 1) also resets the record;
 * rg_ctx_kernel_time returns the number of timed launches and their summed
 * device time since then. */
int rg_ctx_timing(rg_ctx* ctx, int enable);
/* Device stream-copy probe: copies `bytes` between two device buffers `reps`
 * times with a plain 16-byte-per-lane kernel (k_stream_copy); *gbps counts
 * read + write bytes.  The measured HBM reference beside the 8 TB/s peak,
 * and the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE
 * counters.  No counterpart in the reference. */
int rg_device_copy_bandwidth(rg_ctx* ctx, int64_t bytes, int reps, double* gbps, double* ms_per_copy);
/* Device read probes, the read-side references for the FETCH_SIZE counter:
 * k_stream_read loads `bytes` once per pass (16 bytes per lane, nothing
 * stored), *gbps = bytes / time; then, when gathers > 0, k_gather_read makes
 * `gathers` random 8-byte loads over the same buffer per pass
 * (*ms_per_gather_pass).  `reps` passes of each are timed.  No counterpart
 * in the reference. */
int rg_device_read_bandwidth(rg_ctx* ctx, int64_t bytes, int reps, int64_t gathers, double* gbps,
                             double* ms_per_read, double* ms_per_gather_pass);
int rg_ctx_kernel_time(rg_ctx* ctx, int64_t* launches, double* total_ms);
/* Iteration loops as HIP graphs.  enable is a set of flags:
 * RG_GRAPHS_LOOPS (on by default): the blocked PageRank engine captures its
 * whole iteration loop once per context and replays it with one launch
 * while the graph, the iteration count and the context's buffers stay the
 * same (a serving slot answering repeated PageRank queries), and a
 * bit-parallel BFS batch replays a captured chunk of four levels; captured
 * again when they change.  Not used while per-launch timing (rg_ctx_timing)
 * is on.  RG_GRAPHS_BFS_LEVELS (off by default): the single-source pull BFS
 * decides each level's kind (top-down or bottom-up) on the device and runs
 * chunks of four levels (replayed as a captured graph with RG_GRAPHS_LOOPS)
 * with one host check per chunk instead of a read of the frontier counts
 * per level -- same answers; on RMAT-26 the launches of the kernels a level
 * does not use cost what the per-level reads cost (2.90 against 2.88 ms
 * per query).  rg_ctx_graphs(ctx, 0) launches every kernel of every
 * iteration from the host.  No counterpart in the reference. */
enum { RG_GRAPHS_LOOPS = 1, RG_GRAPHS_BFS_LEVELS = 2 };
int rg_ctx_graphs(rg_ctx* ctx, int enable);

/* ---- device graphs ------------------------------------------------------- */
typedef struct rg_graph rg_graph;
/* Upload a host graph: uint64 offsets, uint32 neighbours, plus the CSC for a
 * directed graph (Ligra asymmetricVertex keeps both, vertex.h:244-362). */
int rg_graph_upload(rg_ctx* ctx, const rg_hgraph* hg, rg_graph** out);
int rg_graph_info(const rg_graph* g, int64_t* n, int64_t* m, int* symmetric);
/* Copy the current orientation's CSR (in_edges=0) or CSC back to the host.
 * offsets has n+1 entries, nbrs m entries. */
int rg_graph_download(const rg_graph* g, int in_edges, uint64_t* offsets, uint32_t* nbrs);
/* Fast ingest of the binary formats, file -> device with no host CSR
 * (replaces Ligra's binary reader, PNO Ligra/ligra/IO.h:477-626, and GAP's
 * ReadSerializedGraph, GAP/src/reader.h:251-304, for callers whose graph
 * lives on the device): the file is memory-mapped and streamed through two
 * pinned staging buffers (host copies overlap the DMA).  format RG_FMT_SG
 * (the file says directed or not; a directed file's in-edge CSR is read
 * as stored) or RG_FMT_ADJ_BINARY (path = the prefix of .config/.adj/.idx;
 * symmetric as Ligra's -s; a directed graph's in-edges are built on the
 * device, lists sorted by source, as the host reader makes them).  The
 * lists are taken as stored (no dedup, as the binary readers).  A loaded
 * CSR that is not valid (offsets not starting at 0, decreasing, not ending
 * at m; a neighbour >= n) is refused with RG_EFORMAT before any kernel uses
 * it.  When <path>.newids exists (written by rg_graph_save), the graph is
 * marked degree-sorted and new_ids (host, n entries, nullable) receives it.
 * No counterpart in the reference beyond the readers it replaces. */
int rg_graph_load(rg_ctx* ctx, const char* path, int format, int symmetric, uint32_t* new_ids, rg_graph** out);
/* n, m, symmetric (-1 for Ligra binary, whose files do not say) and whether
 * a new_ids sidecar exists, from the file headers. */
int rg_graph_file_info(const char* path, int format, int64_t* n, int64_t* m, int* symmetric, int* has_new_ids);
/* Writes a device graph in its original orientation as RG_FMT_SG (GAP
 * WriteSerializedGraph, GAP/src/writer.h:39-69) or RG_FMT_ADJ_BINARY
 * (fewer than 2^32 arcs), streamed through pinned buffers; with new_ids
 * (host, n entries) also <path>.newids, n uint32 new_ids[orig] -- the cache
 * of a DegreeSorted graph that a restart loads with rg_graph_load instead of
 * building and sorting the graph again.  Without new_ids a stale sidecar is
 * removed. */
int rg_graph_save(const rg_graph* g, const char* path, int format, const uint32_t* new_ids);
/* Ligra graph::transpose() (graph.h:78-86): swap CSR and CSC of a directed
 * graph; no-op for symmetric graphs. */
int rg_graph_transpose(rg_graph* g);
int rg_graph_free(rg_graph* g);

/* DegreeSorting on the GPU (PNO/DS Ligra/ligra/IO.h:729-944): order vertices
 * by (degree, id) descending -- std::greater on (degree, id) pairs, IO.h:755 --
 * and rebuild CSR (+CSC) with relabelled neighbours.  by_outdeg=1 sorts by
 * out-degree (-preprocess 0), 0 by in-degree (-preprocess 1; symmetric graphs
 * use their degree either way).  sort_neighbors=1 is the strict PNO copy
 * (lists sorted), 0 the relaxed DS/RD copy (IO.h:725-726: original order,
 * relabelled).  new_ids (host, n entries, nullable) receives new_ids[orig].
 * The result is a new device graph; `g` is left unchanged.  Graphs of fewer
 * than 2^31 vertices are accepted at any arc count (strict lists of graphs
 * with 2^31 arcs or more are sorted in chunks of at most 2^30 arcs). */
int rg_graph_degree_sort(rg_graph* g, int by_outdeg, int sort_neighbors,
                         rg_graph** out, uint32_t* new_ids);

/* ---- apps ------------------------------------------------------------------
 * Variants (the five reference trees):
 *   RG_PUSH     push-only edgeMap, dense-forward (PNO; also DS on a sorted graph)
 *   RG_PULL     Ligra direction switching with pull dense rounds (PP)
 *   RG_HUBDUP   push with private hub copies + merge (HD), hubs = top-k
 *               in-degree vertices found by a degree threshold
 *   RG_RADAR    HUBDUP with hubs = ids 0..k-1 (RD, expects a degree-sorted graph)
 *   RG_BLOCKED  PageRank only: the all-vertices push without global atomics --
 *               destinations in LDS-resident blocks (hub blocks split over
 *               several private copies merged by the vertex phase), in-edges
 *               ordered by source segment so reads stay in each XCD's L2.
 *               On a degree-sorted graph this is RADAR's reorder + duplicate
 *               scheme mapped onto LDS and L2.
 * Hub placement: RG_HUBS_LDS (one copy per workgroup in LDS) or
 * RG_HUBS_L2 (one copy per XCD in global memory, L2-resident).
 */
enum { RG_PUSH = 0, RG_PULL = 1, RG_HUBDUP = 2, RG_RADAR = 3, RG_BLOCKED = 4 };
enum { RG_HUBS_LDS = 0, RG_HUBS_L2 = 1 };

typedef struct rg_opts {
  int variant;          /* RG_PUSH ...                                        */
  int hub_placement;    /* RG_HUBS_LDS / RG_HUBS_L2                           */
  int64_t hubs;         /* k of the LDS tier; <= 0: what leaves room for one  */
                        /* workgroup per CU (PageRank, BC backward) or four   */
                        /* (BFS, Radii, BC forward: latency-bound)            */
  int64_t threshold;    /* sparse/dense switch; -1 = m/20 (ligra.h:269)       */
  int wg_size;          /* workgroup size for edge kernels; 0 = default       */
  int heavy_degree;     /* wave-per-vertex cut-over degree; 0 = default       */
  int64_t l2_hubs;      /* HUBDUP/RADAR PageRank and BC backward: hub slots    */
                        /* [hubs, l2_hubs) get one copy per XCD; <= hubs = off */
                        /* (the default: measured no gain)                     */
  int dst_block_bits;   /* RG_BLOCKED: 2^bits destinations per LDS block       */
                        /* (10..14; 0 = 14)                                    */
  int src_segment_bits; /* RG_BLOCKED: 2^bits sources per L2 segment (1..24;   */
                        /* 0 = order by source id)                             */
  int unroll;           /* RG_BLOCKED: edges in flight per lane (1 or 4; 0 =   */
                        /* 4 on a graph made by rg_graph_degree_sort, else 1) */
  int64_t hub_min_arcs; /* HUBDUP/RADAR: list frontiers with fewer arcs        */
                        /* (|F| + out-degree sum) run without hub copies; 0 =  */
                        /* 4096 per CU, < 0 = always with copies               */
  int unit_order;       /* RG_BLOCKED: order in which units run: 0 = the split */
                        /* blocks' units by first source, then the single-    */
                        /* unit blocks; 1 = that order cut into per-XCD       */
                        /* shares (a queue per XCD); 2 = the layout's order   */
  int unit_queue;       /* RG_BLOCKED: 0 = persistent workgroups take units   */
                        /* from a work queue; 1 = one workgroup per unit      */
  int batch_launch;     /* bit-parallel batches (rg_bfs_batch, rg_radii's     */
                        /* batch, serving BFS batches): 1 = per-level kernels */
                        /* (chunks of four levels replayed as HIP graphs, one */
                        /* host check per chunk); 2 = the whole batch as one  */
                        /* persistent cooperative launch (grid barriers       */
                        /* between phases); 0 = the library's choice.         */
                        /* rg_bc_batch: 0 = per-level kernels, a last batch   */
                        /* of fewer than 12 sources as single-source runs;    */
                        /* 1 = per-level kernels for every batch; 2 = each    */
                        /* batch (forward and backward) as one persistent     */
                        /* launch                                             */
} rg_opts;

/* Fill `o` with the defaults (RG_PUSH, threshold -1, library-chosen k). */
void rg_opts_default(rg_opts* o);

/* PageRank (PNO Ligra/apps/PageRank.C:92-141; HD PageRank.C; RD PageRank.C):
 * ranks start at 1/n; per iteration p_next[d] += p_curr[s]/outdeg(s) over all
 * arcs, p_next = 0.85 p_next + 0.15/n, L1 = sum |p_curr - p_next|, stop when
 * L1 < 1e-7 or after maxiters.  iters_out gets the reference's printed
 * "Num Iters" (maxiters+1 when not converged), l1_out the last L1, ranks_out
 * (host, n, nullable) the last computed ranks.  ms_out (nullable) gets the
 * device time of the iteration loop in milliseconds. */
int rg_pagerank(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                int64_t* iters_out, double* l1_out, double* ms_out);

/* PageRankDelta (PNO Ligra/apps/PageRankDelta.C:102-154): p = 0, Delta =
 * 1/n, every vertex active; per round nghSum[d] += Delta[s]/outdeg(s) over
 * the arcs of active s (edgeMap with Ligra's m/20 switch), then the
 * vertexFilter (round 1: Delta = 0.85 nghSum + 0.15/n, p += Delta, Delta -=
 * 1/n; later rounds: Delta = 0.85 nghSum, p += Delta when kept) keeps the
 * vertices with |Delta| > 0.01 p; stop when sum |Delta| < 1e-7 or after
 * maxiters.  iters_out = printed "Num Iters" (maxiters+1 when not
 * converged), l1_out = "L1_Norm", ranks_out (nullable) = p.  RG_BLOCKED
 * runs dense rounds on the blocked PageRank layout, each as one edge pass
 * with the vertexFilter fused (the active list comes out in no fixed
 * order); sparse rounds push. */
int rg_pagerank_delta(rg_graph* g, int64_t maxiters, const rg_opts* o, double* ranks_out,
                      int64_t* iters_out, double* l1_out, double* ms_out);

/* BFS (PNO Ligra/apps/BFS.C:73-117): parents via first successful CAS,
 * parents[src] = src, UINT32_MAX when unreached.  iters_out = rounds
 * (printed "Num. Iters"), reached_out = tree size, edges_out = sum of
 * out-degrees of reached vertices (the traversed-edge count for TEPS). */
int rg_bfs(rg_graph* g, uint32_t src, const rg_opts* o, uint32_t* parents_out,
           int64_t* iters_out, int64_t* reached_out, int64_t* edges_out, double* ms_out);

/* Summary of one BFS query: iters = rounds with a nonempty frontier (the
 * printed "Num. Iters", BFS.C:107), reached = tree size (BFS.C:108-115),
 * edges = sum of out-degrees of the reached vertices (TEPS numerator),
 * checksum = sum over reached v of rg_level_hash(v, level(v)) mod 2^64, which
 * identifies the level array.  No counterpart in the reference beyond the
 * first two (its BFS prints only those). */
typedef struct rg_bfs_summary {
  int64_t iters, reached, edges;
  uint64_t checksum;
} rg_bfs_summary;
uint64_t rg_level_hash(uint32_t v, uint32_t level);

/* One BFS query on the single-source engine (BFS.C:73-117 semantics, as
 * rg_bfs), returning its summary. */
int rg_bfs_query(rg_graph* g, uint32_t src, const rg_opts* o, rg_bfs_summary* out, double* ms_out);

/* Batched BFS: 1..256 sources traversed together as a bit-parallel frontier
 * (source i owns bit i of ceil(n_src/64) words per vertex; Radii's trick,
 * Radii.C:35-75, generalised).  out[i] equals rg_bfs_query(sources[i]) for
 * every i; levels_out (nullable, nsources * n int32, source-major) receives
 * each source's level array (-1 = unreached).  o: threshold (sparse/dense
 * level switch, < 0 = m/20), heavy_degree, wg_size. */
int rg_bfs_batch(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, int32_t* levels_out,
                 rg_bfs_summary* out, double* ms_out);

/* Radii (PNO Ligra/apps/Radii.C:98-157): bit-parallel BFS from up to 64
 * sources (source i owns bit i); radii[v] = last round in which v's visited
 * word changed, -1 if unreached.  rounds_out = printed "Iters".  With
 * o->variant RG_RADAR or RG_BLOCKED the starts run as one batch of the
 * bit-parallel BFS (rg_bfs_batch's engine: blocked dense levels). */
int rg_radii(rg_graph* g, const uint32_t* starts, int nstarts, const rg_opts* o,
             int32_t* radii_out, int64_t* rounds_out, double* ms_out);

/* BC (PNO Ligra/apps/BC-manyiter-v2.C:124-241), sources given in the graph's
 * own id space, processed in order with the reference's semantics: forward
 * path counting by level, backward dependency accumulation on the transposed
 * graph (the transpose is not undone between sources), Dependencies carried
 * over between sources (zeroed once), final (dep - 1/sigma)/(1/sigma).
 * combined_out = sum of forward rounds (printed "Combined Iters").  With
 * o->variant RG_RADAR or RG_BLOCKED, levels above the sparse/dense threshold
 * run on the destination-blocked layout (LDS sums, no global atomics). */
int rg_bc(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o,
          double* deps_out, int64_t* combined_out, double* ms_out);

/* Summary of one single-source BC query: rounds = forward rounds (the
 * "Combined Iters" of rg_bc with that one source), reached = vertices the
 * source reaches, wsum = sum over the finite dependencies dep[v] of
 * dep[v] * w(v) with w(v) = (rg_level_hash(v, 0xffffffff) >> 11) * 2^-53,
 * nonfinite = count of non-finite dependencies (the unreached vertices are
 * NaN in BC-manyiter-v2.C:230-232's final step).  No counterpart in the
 * reference, which prints no per-source result. */
typedef struct rg_bc_summary {
  int64_t rounds, reached;
  double wsum;
  int64_t nonfinite;
} rg_bc_summary;

/* Batched BC: 1..256 independent single-source queries, answered in
 * batches of 64 (in order; the last batch takes the remainder); row i of
 * deps_out (nullable, nsources * n doubles, source-major) equals
 * rg_bc(g, &sources[i], 1, ...)'s deps_out within FP64 rounding, out[i] its
 * summary.  A batch's forward levels run on the bit-parallel BFS of
 * rg_bfs_batch (source i owns bit i); path counts and dependencies are
 * summed per vertex by one wavefront with a lane per source, with no
 * atomics but for vertices of degree above 4096.  A batch with a source
 * deeper than 63 levels is answered by single-source runs, and so is a last
 * batch of fewer than 12 sources when o->batch_launch is 0 (such a small
 * batch costs more than its single-source runs).  o->batch_launch 2 runs
 * each batch, forward and backward, as one persistent launch.  No
 * counterpart in the reference. */
int rg_bc_batch(rg_graph* g, const uint32_t* sources, int nsources, const rg_opts* o, double* deps_out,
                rg_bc_summary* out, double* ms_out);

/* Triangle counting (PNO GAP/src/tc_local.cc:48-86): u > v > w ordered merge
 * counting over a symmetric graph with sorted lists; total and per-vertex
 * counts (each triangle adds 1 to all three corners).  per_vertex nullable.
 * Totals only: the leading vertices of degree >= o->heavy_degree (0 or a
 * null o: 512; at most the first 2^19 ids) are counted from their lower
 * side with an LDS bitmap, the rest per upper vertex. */
int rg_tc(rg_graph* g, const rg_opts* o, uint64_t* total_out, uint64_t* per_vertex,
          double* ms_out);

/* ---- engine level: Ligra's vertexSubset / edgeMap / vertexMap -----------------
 * The reference's operator interface (PNO Ligra/ligra/ligra.h:336-340
 * edgeMap, 447-479 vertexMap; vertexSubset.h:113-229) with each app functor
 * as a named operator over device state arrays.  A subset is a device
 * vertex list (or "all vertices"); edgeMap chooses sparse / dense rounds
 * with Ligra's rule |F| + sum(outdeg) > threshold (threshold < 0: m/20,
 * ligra.h:269) and returns the next frontier unless RG_NO_OUTPUT. */
typedef struct rg_subset rg_subset;
typedef struct rg_array rg_array;
enum { RG_NO_OUTPUT = 1, RG_DENSE_FORWARD = 8 };   /* Ligra's flags (ligra.h:51-58) */
enum { RG_OP_BFS = 0, RG_OP_PAGERANK = 1, RG_OP_RADII = 2, RG_OP_BC_FWD = 3, RG_OP_BC_BACK = 4 };
enum { RG_VOP_PR_UPDATE = 0, RG_VOP_PR_CONTRIB = 1, RG_VOP_RADII_SYNC = 2, RG_VOP_MARK = 3,
       RG_VOP_BC_BACK = 4, RG_VOP_INVERT = 5, RG_VOP_BC_FINAL = 6, RG_VOP_ABS_DIFF = 7, RG_VOP_ABS = 8 };
enum { RG_FOP_PRD_FIRST = 0, RG_FOP_PRD = 1, RG_FOP_U32_EQ = 2, RG_FOP_U8_NONZERO = 3 };
int rg_subset_create(rg_graph* g, rg_subset** out);                 /* empty subset  */
int rg_subset_set(rg_subset* s, const uint32_t* ids, int64_t count); /* from host ids */
int rg_subset_set_all(rg_subset* s);                                 /* every vertex  */
int rg_subset_size(const rg_subset* s, int64_t* size);               /* numNonzeros   */
int rg_subset_get(const rg_subset* s, uint32_t* ids);                /* toSparse copy */
int rg_subset_free(rg_subset* s);                                    /* del()         */
/* Device state arrays (the apps' malloc'd arrays): elem_bytes 1, 4 or 8. */
int rg_array_alloc(rg_ctx* ctx, int64_t count, int elem_bytes, rg_array** out);
int rg_array_write(rg_array* a, const void* host, int64_t first, int64_t count);
int rg_array_read(const rg_array* a, void* host, int64_t first, int64_t count);
int rg_array_fill(rg_array* a, uint64_t bits);  /* every element = low elem_bytes of bits */
int rg_array_free(rg_array* a);
/* sequence::plusReduce (PNO Ligra/ligra/sequence.h, as PageRank.C:128 and
 * PageRankDelta.C:138 call it) over an array of doubles, summed in a fixed
 * order (the same bits on every call). */
int rg_array_sum(const rg_array* a, double* out);
/* The array's device address, for entry points that take device pointers
 * into caller memory, such as rg_bfs_noalloc.  No counterpart in the reference. */
int rg_array_device_ptr(const rg_array* a, void** out);
/* edgeMap.  State per operator:
 *   RG_OP_BFS      {parents u32}                    BFS_F      (BFS.C:28-49)
 *   RG_OP_PAGERANK {contrib f64, p_next f64}        PR_F       (PageRank.C:30-43)
 *   RG_OP_RADII    {visited u64, next u64, radii i32} Radii_F  (Radii.C:35-64; `round`)
 *   RG_OP_BC_FWD   {num_paths f64, visited u8}      BC_F       (BC-manyiter-v2.C:32-56)
 *   RG_OP_BC_BACK  {dependencies f64, visited u8}   BC_Back_F  (BC-manyiter-v2.C:58-79)
 * o (nullable) selects push / HUBDUP / RADAR hub copies. */
int rg_edgemap(rg_graph* g, int op, rg_subset* in, rg_array* const* state, int nstate, int64_t round,
               int64_t threshold, int flags, const rg_opts* o, rg_subset* out);
/* vertexMap over a subset:
 *   RG_VOP_PR_UPDATE  {p}: p = a*p + b             PR_Vertex_F (PageRank.C:46-58)
 *   RG_VOP_PR_CONTRIB {p, contrib}: contrib = p/outdeg, every vertex
 *   RG_VOP_RADII_SYNC {visited, next}: visited = next   Radii_Vertex_F (Radii.C:67-75)
 *   RG_VOP_MARK       {visited u8}: visited = 1     BC_Vertex_F (BC-manyiter-v2.C:82-89)
 *   RG_VOP_BC_BACK    {visited u8, deps, inv}: visited = 1, deps += inv
 *                                                   BC_Back_Vertex_F (BC-manyiter-v2.C:93-101)
 *   RG_VOP_INVERT     {x}: x = 1/x, every vertex    (BC-manyiter-v2.C:206-208)
 *   RG_VOP_BC_FINAL   {deps, inv}: deps = (deps - inv)/inv, every vertex (BC-manyiter-v2.C:230-232)
 *   RG_VOP_ABS_DIFF   {x, y}: x = |x - y|           (PageRank.C:125-127, the L1 terms)
 *   RG_VOP_ABS        {x, y}: x = |y|               (PageRankDelta.C:136-137) */
int rg_vertexmap(rg_graph* g, int vop, rg_subset* s, rg_array* const* state, int nstate, double a, double b);
/* vertexFilter(V, filter) (PNO Ligra/ligra/ligra.h:500-509): applies the
 * filter functor to every vertex of s and returns in `out` (which must not
 * be a listed s) the vertices it kept, as an ascending list -- the order of
 * Ligra's dense result.  Functors and their state / params:
 *   RG_FOP_PRD_FIRST  {p, Delta, nghSum} f64; {damping, one_over_n, epsilon2}
 *                     PR_Vertex_F_FirstRound (PageRankDelta.C:49-63)
 *   RG_FOP_PRD        {p, Delta, nghSum} f64; {damping, epsilon2}
 *                     PR_Vertex_F (PageRankDelta.C:65-77)
 *   RG_FOP_U32_EQ     {x u32}; {value}: keeps x[v] == value (e.g. unvisited
 *                     BFS parents, BFS.C:47's commented cond)
 *   RG_FOP_U8_NONZERO {x u8}; {}: keeps x[v] != 0 */
int rg_vertexfilter(rg_graph* g, int fop, rg_subset* s, rg_array* const* state, int nstate, const double* params,
                    int nparams, rg_subset* out);

/* Single-source BFS that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The answer of rg_bfs with the
 * push engine (o->variant RG_PUSH): parents via the first successful CAS
 * into the caller's device array parents_dev (n uint32; the address of an
 * rg_array of 4-byte elements, rg_array_device_ptr, or any device
 * allocation), and out = rg_bfs_query's summary.  The frontier lists, byte
 * map, heavy queue and counters live in one fixed device array of
 * RG_FIXED_SCRATCH_BYTES per GPU, set up when the first context on the GPU
 * is created and held until the process ends; a graph needing more (9 bytes
 * per vertex plus 8 per 683 arcs, about 7M vertices at 16 arcs per vertex)
 * is refused with RG_EINVAL.  Every call on a GPU shares that GPU's array,
 * so calls that overlap on one GPU (other contexts or host threads) race;
 * per-launch timing (rg_ctx_timing) must be off. */
#define RG_FIXED_SCRATCH_BYTES (64u << 20)
int rg_bfs_noalloc(rg_graph* g, uint32_t src, uint32_t* parents_dev, rg_bfs_summary* out);

/* ---- query serving ---------------------------------------------------------
 * A request loop over a resident graph (no counterpart in the reference,
 * whose apps run one Compute per process): queries arrive at given times,
 * `slots` device contexts (one stream each, sharing the graph's arrays)
 * execute them concurrently, and with batching on, every free slot takes
 * all BFS queries that have arrived (up to max_batch) as one bit-parallel
 * batch (rg_bfs_batch) -- continuous batching: queries that arrive while
 * others run join the next batch on the next free slot, once a batch of
 * the waiting queries ends no later than as many single-source runs would
 * (a cost curve timed before the trace's clock starts: one single-source
 * run and batches of 1, 8, 64 and 256 sources, interpolated); otherwise the
 * oldest runs alone on the single-source engine with bfs_opts.  BC queries
 * the same way in batches of up to 64 (rg_bc_batch).  PageRank queries run
 * one per slot.  Admission by device memory: the slots' query state may
 * hold free HBM less hbm_reserve (a configuration where one slot cannot hold
 * a full-width BFS batch plus a BC or PageRank query is refused); a work
 * item starts only when what every slot holds, plus what the items in
 * flight and this one may add, fits -- otherwise a narrower batch, or batch
 * state that is not in use is freed, or the item waits until memory frees
 * up (stats: peak_bytes, mem_waits).  one_at_a_time = 1 runs the same trace on one
 * slot, one query at a time, BFS on the single-source engine
 * (rg_bfs_query with bfs_opts), BC on rg_bc with one source: the baseline.
 * Answers: BFS an rg_bfs_summary (equal to rg_bfs_query's); BC (one source,
 * BC-manyiter-v2.C semantics for that source, orientation restored after)
 * and PageRank (`iters` iterations, PageRank.C) a summary of the result
 * array: wsum = sum over finite x[v] of x[v] * w(v), w(v) =
 * (rg_level_hash(v, 0xffffffff) >> 11) * 2^-53, and the count of
 * non-finite entries (rg_bc_summary's fields). */
enum { RG_Q_BFS = 0, RG_Q_BC = 1, RG_Q_PAGERANK = 2 };
typedef struct rg_query {
  int kind;
  uint32_t source;       /* BFS, BC */
  int32_t iters;         /* PageRank (maxiters) */
  double arrival_ms;     /* offset from the start of the run */
} rg_query;
typedef struct rg_query_result {
  int kind, slot;
  int64_t batch_id, batch_size;
  rg_bfs_summary bfs;
  int64_t iters;         /* PageRank: printed iterations; BC: combined rounds */
  double l1;             /* PageRank: last L1 */
  double wsum;
  int64_t nonfinite;
  double arrival_ms, start_ms, end_ms, latency_ms;
} rg_query_result;
typedef struct rg_serve_config {
  int slots;             /* concurrent device contexts (<= 0: 2) */
  int max_batch;         /* BFS queries per batch (1..256, <= 0: 256) */
  int one_at_a_time;
  int64_t hbm_reserve;   /* bytes of device memory left free (<= 0: 4 GiB) */
  rg_opts opts;          /* engine options for BC, PageRank and batched BFS (default
                            variant RG_BLOCKED: blocked PageRank, blocked dense BC levels) */
  rg_opts bfs_opts;      /* single-source BFS queries (default variant RG_PULL:
                            bottom-up dense levels) */
  int alone_kinds;       /* bit (1 << kind): queries of that kind run alone -- the
                            other slots finish their items and start nothing new
                            until it ends (bandwidth-bound queries run twice as
                            long side by side); default 1 << RG_Q_PAGERANK */
} rg_serve_config;
typedef struct rg_serve_stats {
  int64_t queries, batches;
  int slots;
  double makespan_ms, qps, p50_ms, p99_ms, mean_ms;
  double mean_batch;     /* BFS queries per BFS work item (batch or single run) */
  int64_t hbm_budget, slot_bytes;
  int64_t peak_bytes;    /* most device bytes the slots' query state held (measured + admitted) */
  int64_t mem_waits;     /* work items that waited for device memory, or ran as a narrower batch */
  double mean_bc_batch;  /* BC queries per BC work item */
} rg_serve_stats;
void rg_serve_config_default(rg_serve_config* c);
int rg_serve(rg_graph* g, const rg_serve_config* c, const rg_query* queries, int64_t nqueries,
             rg_query_result* results, rg_serve_stats* stats);

/* ---- autotuner -------------------------------------------------------------
 * Times candidate configurations of `app` on g and writes the fastest to
 * *best (the reference fixes these: llcCap = 1<<25 for the hub count, hubdup
 * Ligra/ligra/IO.h:749-755, and the m/20 threshold, ligra.h:269).
 * RG_APP_PAGERANK: blocked-engine workgroup shape / unroll vs atomic RADAR.
 * RG_APP_BFS: sparse/dense threshold x heavy-degree cut-over x hub count,
 * timed over the given sources.  best_ms (nullable): the winner's time. */
enum { RG_APP_PAGERANK = 0, RG_APP_BFS = 1 };
int rg_autotune(rg_graph* g, int app, const uint32_t* sources, int nsources, const rg_opts* base,
                rg_opts* best, double* best_ms);

#ifdef __cplusplus
}
#endif
#endif /* RADAR_GRAPH_H */
