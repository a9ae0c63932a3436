// PageRank with Ligra's interface (push-new-orig Ligra/apps/PageRank.C:92-141).
// Each iteration is the reference's sequence on the device engine: the
// quotients p_curr/outdeg, edgeMap PR_F over all vertices, vertexMap
// PR_Vertex_F, the L1 terms |p_curr - p_next| summed by plusReduce, the
// reset and the swap (PageRank.C:120-133).  -fused (implied by -variant
// blocked or pull) runs the whole loop as one device loop (rg_pagerank)
// instead.  Usage: PageRank [-s] [-b] [-maxiters K] [-rounds R]
// [-preprocess 0|1] [-variant push|hubdup|radar|blocked|pull] [-k hubs]
// [-fused] [-out file] <graph>
#include <cstdio>

#include "ligra_rg.h"

using namespace ligra_rg;

template <class vertex>
void ligra_rg::Compute(graph<vertex>& GA, commandLine P, pvector<uintE>& new_ids) {
  (void)new_ids;
  Timer tm;
  tm.Start();
  long maxIters = P.getOptionLongValue("-maxiters", 100);
  const long n = GA.n;
  const double damping = 0.85, epsilon = 0.0000001;
  std::vector<double> ranks;
  long iter = 0;
  double L1_norm = 0.0;
  const int variant = engine_opts().variant;
  if (P.getOption("-fused") || variant == RG_BLOCKED || variant == RG_PULL) {
    ranks.resize(n);
    int64_t iters = 0;
    double ms = 0.0;
    check(rg_pagerank(GA.g, maxIters, &engine_opts(), ranks.data(), &iters, &L1_norm, &ms), "PageRank");
    iter = (long)iters;
  } else {
    const double one_over_n = 1 / (double)n;
    array<double> p_curr(GA.ctx, n), p_next(GA.ctx, n), contrib(GA.ctx, n);
    p_curr.fill(one_over_n);
    p_next.fill(0.0);
    vertexSubset Frontier = all_vertices(GA);
    while (iter++ < maxIters) {
      vertexMap(GA, Frontier, PR_Vertex_Contrib_F{p_curr, contrib});
      edgeMap(GA, Frontier, nullptr, PR_F{contrib, p_next}, -1, no_output | dense_forward);
      vertexMap(GA, Frontier, PR_Vertex_F{p_next, damping, 0.0}, damping, (1 - damping) * (1 / (double)n));
      vertexMap(GA, Frontier, Abs_Diff_F{p_curr, p_next});
      L1_norm = plusReduce(p_curr);
      if (L1_norm < epsilon) break;
      p_curr.fill(0.0);   // PR_Vertex_Reset
      swap(p_curr, p_next);
    }
    // The last computed ranks: p_next after a converged iteration, p_curr
    // after the final swap.
    ranks = (iter > maxIters ? p_curr : p_next).get();
    Frontier.del();
    p_curr.del();
    p_next.del();
    contrib.del();
  }
  rg_ctx_sync(GA.ctx);
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Num Iters = " << iter << std::endl;
  std::cout << "[OUTPUT] L1_Norm   = " << L1_norm << std::endl;
  if (char* out = P.getOptionValue("-out")) {
    FILE* fo = std::fopen(out, "wb");
    if (fo) {
      std::fwrite(ranks.data(), sizeof(double), ranks.size(), fo);
      std::fclose(fo);
    }
  }
}

int main(int argc, char** argv) { return ligra_main(argc, argv); }
