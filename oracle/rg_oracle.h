/*
 * rg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's algorithms (CMUAbstract/RADAR-Graph),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker.  Never linked into or called by the product library.
 *
 * Graphs are CSR: uint64 off[n+1], uint32 nbr[m]; directed graphs also pass
 * the in-edge CSR (in_off/in_nbr).  The sequential functions visit vertices
 * and frontiers in ascending id order, which is the order a one-thread run of
 * the reference takes (its OpenMP loops and frontier lists are ascending), so
 * FP64 results match such a run bit-for-bit.
 */
#ifndef RG_ORACLE_H
#define RG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Textbook queue BFS: level[v] = hop distance from src, -1 if unreached.
 * Returns the eccentricity of src (max level). */
int64_t orc_bfs_levels(int64_t n, const uint64_t* off, const uint32_t* nbr, uint32_t src,
                       int32_t* level);

/* Checks a BFS parent array against levels: parent[src]==src, every reached
 * v != src has parent p with an arc p->v and level[p] == level[v]-1, every
 * unreached v has parent UINT32_MAX.  Returns 0 if valid, else 1 + the first
 * offending vertex. */
int64_t orc_bfs_check_parents(int64_t n, const uint64_t* off, const uint32_t* nbr,
                              uint32_t src, const int32_t* level, const uint32_t* parent);

/* PageRank (push-new-orig Ligra/apps/PageRank.C:92-141), sequential push.
 * ranks: the last computed ranks; *iters the printed "Num Iters"; *l1 the
 * last L1 norm (blocked reduction of sequence::plusReduce). nthreads > 1
 * runs the same loop with OpenMP (dynamic,64) and CAS-loop FP64 adds, as the
 * reference does (for timing only: the sum order then varies). */
void orc_pagerank(int64_t n, const uint64_t* off, const uint32_t* nbr, int64_t maxiters,
                  int nthreads, double* ranks, int64_t* iters, double* l1);

/* PageRankDelta (push-new-orig Ligra/apps/PageRankDelta.C:102-154): final
 * ranks p, printed "Num Iters" and L1_Norm. */
void orc_pagerank_delta(int64_t n, const uint64_t* off, const uint32_t* nbr, int64_t maxiters,
                        double* ranks, int64_t* iters, double* l1);

/* Pull PageRank of the push-pull tree (push-pull Ligra/apps/PageRank.C:122,
 * threshold 0 => edgeMapDense every iteration, ligra.h:61-88): p_next[v] sums
 * p_curr[s]/outdeg(s) over in-neighbours s in in-list order. */
void orc_pagerank_pull(int64_t n, const uint64_t* off, const uint64_t* in_off,
                       const uint32_t* in_nbr, int64_t maxiters, double* ranks,
                       int64_t* iters, double* l1);

/* Radii (push-new-orig Ligra/apps/Radii.C:98-157) from the given starts
 * (at most 64; start i owns bit i).  radii[v] = -1 if unreached. */
void orc_radii(int64_t n, const uint64_t* off, const uint32_t* nbr, const uint32_t* starts,
               int nstarts, int32_t* radii, int64_t* rounds);

/* BC (push-new-orig Ligra/apps/BC-manyiter-v2.C:124-241) over the given
 * sources (graph id space), with the reference's carry-over of
 * Dependencies between sources and its per-source transpose of a directed
 * graph.  For symmetric graphs pass in_off/in_nbr = off/nbr. */
void orc_bc(int64_t n, const uint64_t* off, const uint32_t* nbr, const uint64_t* in_off,
            const uint32_t* in_nbr, const uint32_t* sources, int nsources, double* deps,
            int64_t* combined_iters);

/* Triangle counting (push-new-orig GAP/src/tc_local.cc:48-86): u > v > w with
 * sorted lists; per_vertex (nullable) gets +1 per triangle corner. */
uint64_t orc_tc(int64_t n, const uint64_t* off, const uint32_t* nbr, uint64_t* per_vertex);

/* Brute-force triangle count by sorted-set intersection, the reference's own
 * verifier (tc_local.cc:129-148): sum over arcs (u,v) of |N(u) & N(v)| / 6. */
uint64_t orc_tc_verify(int64_t n, const uint64_t* off, const uint32_t* nbr);

/* DegreeSorting (push-new-orig Ligra/ligra/IO.h:729-944, relaxed copy in
 * degree-sorting/Ligra/ligra/IO.h:729-894).  key_deg[v] is the sort key
 * (in- or out-degree); order by (key, id) descending; new_ids[orig] = rank.
 * Rebuilds the CSR given by (off, nbr) into (noff, nnbr): list of new vertex
 * new_ids[v] = relabelled list of v in original order, sorted when `strict`.
 * Call once for the CSR and once for the CSC of a directed graph. */
void orc_degree_order(int64_t n, const uint32_t* key_deg, uint32_t* new_ids);
void orc_relabel_csr(int64_t n, const uint64_t* off, const uint32_t* nbr,
                     const uint32_t* new_ids, int strict, uint64_t* noff, uint32_t* nnbr);

/* HUBDUP hub selection (hubdup Ligra/ligra/IO.h:726-804): hubs are the
 * vertices whose degree exceeds the (num_hubs)-th largest degree (0-based).
 * is_hub[v] set to 0/1; returns the hub count.  num_hubs must be < n. */
int64_t orc_identify_hubs(int64_t n, const uint32_t* deg, int64_t num_hubs, uint8_t* is_hub);

/* Ligra's hashInt(ulong) (utils.h:347-355). */
uint64_t orc_hash_int(uint64_t a);

#ifdef __cplusplus
}
#endif
#endif
