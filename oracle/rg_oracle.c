/*
 * rg_oracle.c -- TEST INFRASTRUCTURE ONLY (see rg_oracle.h).
 *
 * Each function restates one reference algorithm; citations are relative to
 * the reference's src/ tree.  Loops run in ascending vertex / frontier order,
 * matching a one-thread run of the reference, unless a function says
 * otherwise.
 */
#include "rg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NONE32 0xffffffffu

/* ---- BFS ------------------------------------------------------------------ */

int64_t orc_bfs_levels(int64_t n, const uint64_t* off, const uint32_t* nbr, uint32_t src,
                       int32_t* level) {
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) level[i] = -1;
  int64_t head = 0, tail = 0, ecc = 0;
  level[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
      uint32_t v = nbr[j];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        if (level[v] > ecc) ecc = level[v];
        q[tail++] = v;
      }
    }
  }
  free(q);
  return ecc;
}

int64_t orc_bfs_check_parents(int64_t n, const uint64_t* off, const uint32_t* nbr,
                              uint32_t src, const int32_t* level, const uint32_t* parent) {
  /* Graph500-style validation: the root is its own parent; every reached
   * vertex has a parent one level up with an arc parent -> v; unreached
   * vertices have none.  The arc test is one pass over all arcs (linear in
   * m whatever the degrees or the order of the lists). */
  if (parent[src] != src) return 1 + (int64_t)src;
  for (int64_t v = 0; v < n; ++v) {
    if (v == (int64_t)src) continue;
    uint32_t p = parent[v];
    if (level[v] < 0) {
      if (p != NONE32) return 1 + v;
      continue;
    }
    if (p == NONE32 || p >= (uint64_t)n) return 1 + v;
    if (level[p] != level[v] - 1) return 1 + v;
  }
  unsigned char* ok = (unsigned char*)calloc((size_t)n + 1, 1);
  if (!ok) return -1;
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t j = off[u]; j < off[u + 1]; ++j)
      if (parent[nbr[j]] == (uint32_t)u) ok[nbr[j]] = 1;
  int64_t bad = 0;
  for (int64_t v = 0; v < n && !bad; ++v)
    if (v != (int64_t)src && level[v] >= 0 && !ok[v]) bad = 1 + v;
  free(ok);
  return bad;
}

/* ---- PageRank --------------------------------------------------------------
 * push-new-orig Ligra/apps/PageRank.C:92-141.  sequence::plusReduce
 * (Ligra/ligra/utils.h:133-142, 151-153) sums blocks of 1024 serially and
 * then reduces the block sums the same way, recursively. */

static double blocked_sum(const double* a, int64_t n) {
  if (n <= 0) return 0.0;
  int64_t nb = 1 + (n - 1) / 1024;
  if (nb <= 1) {
    double r = a[0];
    for (int64_t j = 1; j < n; ++j) r = r + a[j];
    return r;
  }
  double* sums = (double*)malloc(sizeof(double) * (size_t)nb);
  for (int64_t b = 0; b < nb; ++b) {
    int64_t s = b * 1024, e = s + 1024 < n ? s + 1024 : n;
    double r = a[s];
    for (int64_t j = s + 1; j < e; ++j) r = r + a[j];
    sums[b] = r;
  }
  double r = blocked_sum(sums, nb);
  free(sums);
  return r;
}

static void cas_add(double* p, double x) {
  /* writeAdd (Ligra/ligra/utils.h:330-335): CAS loop on the 8-byte word. */
  union { double d; long long l; } o, nw;
  do {
    o.d = *(volatile double*)p;
    nw.d = o.d + x;
  } while (!__sync_bool_compare_and_swap((long long*)p, o.l, nw.l));
}

void orc_pagerank(int64_t n, const uint64_t* off, const uint32_t* nbr, int64_t maxiters,
                  int nthreads, double* ranks, int64_t* iters, double* l1) {
  const double damping = 0.85, epsilon = 0.0000001;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * (1 / (double)n);
  double* p_curr = (double*)malloc(sizeof(double) * (size_t)n);
  double* p_next = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) { p_curr[i] = one_over_n; p_next[i] = 0; }
  int64_t iter = 0;
  double L1 = 0.0;
  if (nthreads < 1) nthreads = 1;
  while (iter++ < maxiters) {
    if (nthreads == 1) {
      for (int64_t s = 0; s < n; ++s) {
        uint64_t deg = off[s + 1] - off[s];
        for (uint64_t j = off[s]; j < off[s + 1]; ++j)
          p_next[nbr[j]] += p_curr[s] / (double)(uint32_t)deg;
      }
    } else {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
      for (int64_t s = 0; s < n; ++s) {
        uint64_t deg = off[s + 1] - off[s];
        for (uint64_t j = off[s]; j < off[s + 1]; ++j)
          cas_add(&p_next[nbr[j]], p_curr[s] / (double)(uint32_t)deg);
      }
    }
    for (int64_t i = 0; i < n; ++i) p_next[i] = damping * p_next[i] + added;
    for (int64_t i = 0; i < n; ++i) p_curr[i] = fabs(p_curr[i] - p_next[i]);
    L1 = blocked_sum(p_curr, n);
    if (L1 < epsilon) break;
    for (int64_t i = 0; i < n; ++i) p_curr[i] = 0.0;
    double* t = p_curr; p_curr = p_next; p_next = t;
  }
  /* On break the new ranks are in p_next; otherwise the final swap moved
   * them into p_curr. */
  const double* fin = (iter <= maxiters) ? p_next : p_curr;
  memcpy(ranks, fin, sizeof(double) * (size_t)n);
  *iters = iter;
  *l1 = L1;
  free(p_curr);
  free(p_next);
}

/* ---- PageRankDelta -----------------------------------------------------------
 * push-new-orig Ligra/apps/PageRankDelta.C:102-154: p = 0, Delta = 1/n,
 * every vertex active; per round nghSum[d] += Delta[s]/outdeg(s) over the
 * arcs of active s (PR_F, 30-48), then PR_Vertex_F_FirstRound (51-64) in
 * round 1 and PR_Vertex_F (66-78) afterwards decide the next active set
 * (|Delta| > 0.01 p), L1 = sum |Delta| (plusReduce), stop when L1 < 1e-7.
 * Sources ascending, neighbours in list order: a one-thread run's order. */
void orc_pagerank_delta(int64_t n, const uint64_t* off, const uint32_t* nbr, int64_t maxiters,
                        double* ranks, int64_t* iters, double* l1) {
  const double damping = 0.85, epsilon = 0.0000001, epsilon2 = 0.01;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * one_over_n;
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* delta = (double*)malloc(sizeof(double) * (size_t)n);
  double* ngh = (double*)malloc(sizeof(double) * (size_t)n);
  unsigned char* fr = (unsigned char*)malloc((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    p[i] = 0.0;
    delta[i] = one_over_n;
    ngh[i] = 0.0;
    fr[i] = 1;
  }
  int64_t round = 0;
  double L1 = 0.0;
  while (round++ < maxiters) {
    for (int64_t s = 0; s < n; ++s) {
      if (!fr[s]) continue;
      uint64_t deg = off[s + 1] - off[s];
      for (uint64_t j = off[s]; j < off[s + 1]; ++j) ngh[nbr[j]] = ngh[nbr[j]] + delta[s] / (double)(uint32_t)deg;
    }
    for (int64_t i = 0; i < n; ++i) {
      if (round == 1) {
        delta[i] = damping * ngh[i] + added;
        p[i] += delta[i];
        delta[i] -= one_over_n;
        fr[i] = fabs(delta[i]) > epsilon2 * p[i];
      } else {
        delta[i] = ngh[i] * damping;
        if (fabs(delta[i]) > epsilon2 * p[i]) {
          p[i] += delta[i];
          fr[i] = 1;
        } else {
          fr[i] = 0;
        }
      }
    }
    for (int64_t i = 0; i < n; ++i) ngh[i] = fabs(delta[i]);
    L1 = blocked_sum(ngh, n);
    if (L1 < epsilon) break;
    for (int64_t i = 0; i < n; ++i) ngh[i] = 0.0;
  }
  memcpy(ranks, p, sizeof(double) * (size_t)n);
  *iters = round;
  *l1 = L1;
  free(p);
  free(delta);
  free(ngh);
  free(fr);
}

void orc_pagerank_pull(int64_t n, const uint64_t* off, const uint64_t* in_off,
                       const uint32_t* in_nbr, int64_t maxiters, double* ranks,
                       int64_t* iters, double* l1) {
  const double damping = 0.85, epsilon = 0.0000001;
  const double one_over_n = 1 / (double)n;
  const double added = (1 - damping) * (1 / (double)n);
  double* p_curr = (double*)malloc(sizeof(double) * (size_t)n);
  double* p_next = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) { p_curr[i] = one_over_n; p_next[i] = 0; }
  int64_t iter = 0;
  double L1 = 0.0;
  while (iter++ < maxiters) {
    for (int64_t v = 0; v < n; ++v)
      for (uint64_t j = in_off[v]; j < in_off[v + 1]; ++j) {
        uint32_t s = in_nbr[j];
        p_next[v] += p_curr[s] / (double)(uint32_t)(off[s + 1] - off[s]);
      }
    for (int64_t i = 0; i < n; ++i) p_next[i] = damping * p_next[i] + added;
    for (int64_t i = 0; i < n; ++i) p_curr[i] = fabs(p_curr[i] - p_next[i]);
    L1 = blocked_sum(p_curr, n);
    if (L1 < epsilon) break;
    for (int64_t i = 0; i < n; ++i) p_curr[i] = 0.0;
    double* t = p_curr; p_curr = p_next; p_next = t;
  }
  const double* fin = (iter <= maxiters) ? p_next : p_curr;
  memcpy(ranks, fin, sizeof(double) * (size_t)n);
  *iters = iter;
  *l1 = L1;
  free(p_curr);
  free(p_next);
}

/* ---- Radii -----------------------------------------------------------------
 * push-new-orig Ligra/apps/Radii.C:35-75 (functors), 98-157 (Compute).
 * Each round: Visited = NextVisited on the frontier (vertexMap), then for
 * every frontier s and out-neighbour d: toWrite = V[d] | V[s]; if it adds
 * bits, NextVisited[d] |= toWrite and, the first time in this round,
 * radii[d] = round and d joins the next frontier. */

void orc_radii(int64_t n, const uint64_t* off, const uint32_t* nbr, const uint32_t* starts,
               int nstarts, int32_t* radii, int64_t* rounds) {
  int64_t* vis = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  int64_t* nvis = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  uint8_t* in_next = (uint8_t*)calloc((size_t)n, 1);
  uint32_t* front = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 64 ? n : 64));
  uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 64 ? n : 64));
  for (int64_t i = 0; i < n; ++i) radii[i] = -1;
  int64_t fs = 0;
  for (int i = 0; i < nstarts; ++i) {
    uint32_t v = starts[i];
    radii[v] = 0;
    front[fs++] = v;
    nvis[v] = (int64_t)1 << i;  /* Radii.C:126 assigns, so a repeated start keeps the last bit */
  }
  int32_t round = 0;
  while (fs > 0) {
    round++;
    for (int64_t i = 0; i < fs; ++i) vis[front[i]] = nvis[front[i]];
    int64_t ns = 0;
    for (int64_t i = 0; i < fs; ++i) {
      uint32_t s = front[i];
      for (uint64_t j = off[s]; j < off[s + 1]; ++j) {
        uint32_t d = nbr[j];
        int64_t tw = vis[d] | vis[s];
        if (vis[d] != tw) {
          nvis[d] |= tw;
          if (radii[d] != round) {
            radii[d] = round;
            if (!in_next[d]) { in_next[d] = 1; next[ns++] = d; }
          }
        }
      }
    }
    /* The next frontier is a set; its order does not affect the integers. */
    for (int64_t i = 0; i < ns; ++i) in_next[next[i]] = 0;
    uint32_t* t = front; front = next; next = t;
    fs = ns;
  }
  *rounds = round;
  free(vis); free(nvis); free(in_next); free(front); free(next);
}

/* ---- BC --------------------------------------------------------------------
 * push-new-orig Ligra/apps/BC-manyiter-v2.C:32-101 (functors), 280-398
 * (Compute).  Frontiers are processed in ascending id order (sparse lists
 * from the bitmap scan, ligra.h:150-164, and dense loops are both
 * ascending). */

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

void orc_bc(int64_t n, const uint64_t* off, const uint32_t* nbr, const uint64_t* in_off,
            const uint32_t* in_nbr, const uint32_t* sources, int nsources, double* deps,
            int64_t* combined_iters) {
  double* np_ = (double*)malloc(sizeof(double) * (size_t)n);
  uint8_t* visited = (uint8_t*)malloc((size_t)n);
  uint8_t* mark = (uint8_t*)calloc((size_t)n, 1);
  /* Levels: concatenated frontier lists with start indices. */
  uint32_t* lv = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n + 4);
  int64_t* lstart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 2));
  const uint64_t* co = off;     /* current orientation */
  const uint32_t* cn = nbr;
  const uint64_t* to = in_off;  /* the other one */
  const uint32_t* tn = in_nbr;
  int64_t combined = 0;
  for (int64_t i = 0; i < n; ++i) deps[i] = 0.0;
  for (int it = 0; it < nsources; ++it) {
    uint32_t start = sources[it];
    for (int64_t i = 0; i < n; ++i) { np_[i] = 0.0; visited[i] = 0; }
    np_[start] = 1.0;
    visited[start] = 1;
    int64_t nl = 0;          /* number of levels stored */
    int64_t used = 0;
    lstart[0] = 0;
    lv[used++] = start;
    lstart[++nl] = used;     /* level 0 = {start} */
    int64_t round = 0;
    int64_t fb = 0, fe = 1;  /* current frontier = lv[fb, fe) */
    while (fe > fb) {
      round++;
      int64_t nb = used;
      for (int64_t i = fb; i < fe; ++i) {
        uint32_t s = lv[i];
        for (uint64_t j = co[s]; j < co[s + 1]; ++j) {
          uint32_t d = cn[j];
          if (visited[d] == 0) {            /* BC_F::cond */
            double old = np_[d];
            np_[d] = old + np_[s];          /* BC_F::updateAtomic */
            if (old == 0.0 && !mark[d]) { mark[d] = 1; lv[used++] = d; }
          }
        }
      }
      qsort(lv + nb, (size_t)(used - nb), sizeof(uint32_t), cmp_u32);
      for (int64_t i = nb; i < used; ++i) { mark[lv[i]] = 0; visited[lv[i]] = 1; }
      lstart[++nl] = used;   /* level `round` (possibly empty) */
      fb = nb;
      fe = used;
    }
    combined += round;
    for (int64_t i = 0; i < n; ++i) np_[i] = 1 / np_[i];   /* inverseNumPaths */
    for (int64_t i = 0; i < n; ++i) visited[i] = 0;
    /* Frontier = Levels[round-1]; BC_Back_Vertex_F on it. */
    int64_t L = round - 1;
    for (int64_t i = lstart[L]; i < lstart[L + 1]; ++i) {
      uint32_t v = lv[i];
      visited[v] = 1;
      deps[v] += np_[v];
    }
    /* GA.transpose() -- never undone (BC-manyiter-v2.C:217). */
    { const uint64_t* a = co; co = to; to = a; const uint32_t* b = cn; cn = tn; tn = b; }
    for (int64_t r = round - 2; r >= 0; --r) {
      for (int64_t i = lstart[r + 1]; i < lstart[r + 2]; ++i) {
        uint32_t s = lv[i];
        for (uint64_t j = co[s]; j < co[s + 1]; ++j) {
          uint32_t d = cn[j];
          if (visited[d] == 0) deps[d] += deps[s];   /* BC_Back_F */
        }
      }
      for (int64_t i = lstart[r]; i < lstart[r + 1]; ++i) {
        uint32_t v = lv[i];
        visited[v] = 1;
        deps[v] += np_[v];
      }
    }
    for (int64_t i = 0; i < n; ++i) deps[i] = (deps[i] - np_[i]) / np_[i];
  }
  *combined_iters = combined;
  free(np_); free(visited); free(mark); free(lv); free(lstart);
}

/* ---- Triangle counting ----------------------------------------------------- */

uint64_t orc_tc(int64_t n, const uint64_t* off, const uint32_t* nbr, uint64_t* per_vertex) {
  uint64_t total = 0;
  if (per_vertex) memset(per_vertex, 0, sizeof(uint64_t) * (size_t)n);
  for (int64_t u = 0; u < n; ++u) {
    for (uint64_t a = off[u]; a < off[u + 1]; ++a) {
      uint32_t v = nbr[a];
      if ((int64_t)v > u) break;
      uint64_t it = off[u];
      for (uint64_t b = off[v]; b < off[v + 1]; ++b) {
        uint32_t w = nbr[b];
        if (w > v) break;
        while (nbr[it] < w) it++;
        if (w == nbr[it]) {
          total++;
          if (per_vertex) { per_vertex[u]++; per_vertex[v]++; per_vertex[w]++; }
        }
      }
    }
  }
  return total;
}

uint64_t orc_tc_verify(int64_t n, const uint64_t* off, const uint32_t* nbr) {
  uint64_t total = 0;
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t a = off[u]; a < off[u + 1]; ++a) {
      uint32_t v = nbr[a];
      uint64_t i = off[u], ie = off[u + 1], j = off[v], je = off[v + 1];
      while (i < ie && j < je) {
        if (nbr[i] < nbr[j]) i++;
        else if (nbr[j] < nbr[i]) j++;
        else { total++; i++; j++; }
      }
    }
  return total / 6;
}

/* ---- DegreeSorting ---------------------------------------------------------- */

typedef struct { uint32_t deg, id; } dpair;

static int cmp_dpair_desc(const void* a, const void* b) {
  const dpair* x = (const dpair*)a;
  const dpair* y = (const dpair*)b;
  /* std::greater<pair<uintT,uintE>>: degree descending, then id descending. */
  if (x->deg != y->deg) return x->deg > y->deg ? -1 : 1;
  if (x->id != y->id) return x->id > y->id ? -1 : 1;
  return 0;
}

void orc_degree_order(int64_t n, const uint32_t* key_deg, uint32_t* new_ids) {
  dpair* p = (dpair*)malloc(sizeof(dpair) * (size_t)n);
  for (int64_t v = 0; v < n; ++v) { p[v].deg = key_deg[v]; p[v].id = (uint32_t)v; }
  qsort(p, (size_t)n, sizeof(dpair), cmp_dpair_desc);
  for (int64_t r = 0; r < n; ++r) new_ids[p[r].id] = (uint32_t)r;
  free(p);
}

void orc_relabel_csr(int64_t n, const uint64_t* off, const uint32_t* nbr,
                     const uint32_t* new_ids, int strict, uint64_t* noff, uint32_t* nnbr) {
  uint32_t* inv = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  for (int64_t v = 0; v < n; ++v) inv[new_ids[v]] = (uint32_t)v;
  noff[0] = 0;
  for (int64_t r = 0; r < n; ++r) {
    uint32_t v = inv[r];
    noff[r + 1] = noff[r] + (off[v + 1] - off[v]);
  }
  for (int64_t r = 0; r < n; ++r) {
    uint32_t v = inv[r];
    uint64_t o = noff[r];
    for (uint64_t j = off[v]; j < off[v + 1]; ++j) nnbr[o++] = new_ids[nbr[j]];
    if (strict) qsort(nnbr + noff[r], (size_t)(noff[r + 1] - noff[r]), sizeof(uint32_t), cmp_u32);
  }
  free(inv);
}

/* ---- HUBDUP hub selection ---------------------------------------------------- */

static int cmp_u32_desc(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x < y) - (x > y);
}

int64_t orc_identify_hubs(int64_t n, const uint32_t* deg, int64_t num_hubs, uint8_t* is_hub) {
  /* The reference reads degrees[numHubs] past the array when numHubs >= n
   * (hubdup IO.h:759); here, as in the library, every vertex of positive
   * degree is a hub then. */
  uint32_t thresh = 0;
  if (num_hubs < n) {
    uint32_t* s = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    memcpy(s, deg, sizeof(uint32_t) * (size_t)n);
    qsort(s, (size_t)n, sizeof(uint32_t), cmp_u32_desc);
    thresh = s[num_hubs];            /* hubdup IO.h:759 */
    free(s);
  }
  int64_t c = 0;
  for (int64_t v = 0; v < n; ++v) {
    is_hub[v] = deg[v] > thresh;     /* IO.h:779 */
    c += is_hub[v];
  }
  return c;
}

uint64_t orc_hash_int(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dULL) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddULL) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ULL) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ULL) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cULL) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bULL) ^ (a >> 32);
  return a;
}
