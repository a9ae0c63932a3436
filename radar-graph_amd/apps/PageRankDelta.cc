// PageRankDelta with Ligra's interface (push-new-orig
// Ligra/apps/PageRankDelta.C:90-155).  Each round is the reference's
// sequence on the device engine: edgeMap PR_F over the active vertices
// with the m/20 switch, vertexFilter PR_Vertex_F_FirstRound / PR_Vertex_F
// over all vertices, the L1 norm |Delta| by plusReduce, the reset of
// nghSum, and the filter's output as the next frontier (PageRankDelta.C:
// 128-144).  -fused (implied by -variant blocked) runs the rounds as one
// device loop (rg_pagerank_delta) instead.  Usage: PageRankDelta [-s] [-b]
// [-maxiters K] [-rounds R] [-preprocess 0|1]
// [-variant push|hubdup|radar|blocked] [-k hubs] [-fused] [-out file] <graph>
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
  const double damping = 0.85, epsilon = 0.0000001, epsilon2 = 0.01;
  std::vector<double> ranks;
  long round = 0;
  double L1_norm = 0.0;
  if (P.getOption("-fused") || engine_opts().variant == RG_BLOCKED) {
    ranks.resize(n);
    int64_t iters = 0;
    double ms = 0.0;
    check(rg_pagerank_delta(GA.g, maxIters, &engine_opts(), ranks.data(), &iters, &L1_norm, &ms),
          "PageRankDelta");
    round = (long)iters;
  } else {
    const double one_over_n = 1 / (double)n;
    array<double> p(GA.ctx, n), Delta(GA.ctx, n), nghSum(GA.ctx, n), contrib(GA.ctx, n);
    p.fill(0.0);
    Delta.fill(one_over_n);
    nghSum.fill(0.0);
    vertexSubset Frontier = all_vertices(GA);
    vertexSubset All = all_vertices(GA);
    while (round++ < maxIters) {
      // PR_F's quotient Delta[s]/outdeg(s), then the push over the frontier.
      vertexMap(GA, All, PR_Vertex_Contrib_F{Delta, contrib});
      edgeMap(GA, Frontier, nullptr, PR_F{contrib, nghSum}, GA.m / 20, no_output | dense_forward);
      vertexSubset active =
          (round == 1) ? vertexFilter(GA, All, PR_Vertex_F_FirstRound{p, Delta, nghSum, damping, one_over_n, epsilon2})
                       : vertexFilter(GA, All, PR_Delta_Vertex_F{p, Delta, nghSum, damping, epsilon2});
      // L1 norm, with nghSum as the temporary array.
      vertexMap(GA, All, Abs_F{nghSum, Delta});
      L1_norm = plusReduce(nghSum);
      if (L1_norm < epsilon) {
        active.del();
        break;
      }
      nghSum.fill(0.0);   // PR_Vertex_Reset
      Frontier.del();
      Frontier = active;
    }
    ranks = p.get();
    Frontier.del();
    All.del();
    p.del();
    Delta.del();
    nghSum.del();
    contrib.del();
  }
  rg_ctx_sync(GA.ctx);
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Num Iters = " << round << std::endl;
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
