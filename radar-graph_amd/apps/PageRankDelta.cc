// PageRankDelta with Ligra's interface (push-new-orig
// Ligra/apps/PageRankDelta.C:102-154): the rounds (edgeMap PR_F over the
// active vertices, the vertexFilter, the L1 norm, the reset) run as one
// device loop (rg_pagerank_delta).  Usage: PageRankDelta [-s] [-b]
// [-maxiters K] [-rounds R] [-preprocess 0|1] [-variant push|hubdup|radar|blocked]
// [-k hubs] [-out file] <graph>
#include <cstdio>

#include "ligra_rg.h"

using namespace ligra_rg;

template <class vertex>
void ligra_rg::Compute(graph<vertex>& GA, commandLine P, pvector<uintE>& new_ids) {
  (void)new_ids;
  Timer tm;
  tm.Start();
  long maxIters = P.getOptionLongValue("-maxiters", 100);
  std::vector<double> ranks(GA.n);
  int64_t iters = 0;
  double l1 = 0.0, ms = 0.0;
  check(rg_pagerank_delta(GA.g, maxIters, &engine_opts(), ranks.data(), &iters, &l1, &ms), "PageRankDelta");
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Num Iters = " << iters << std::endl;
  std::cout << "[OUTPUT] L1_Norm   = " << l1 << std::endl;
  if (char* out = P.getOptionValue("-out")) {
    FILE* fo = std::fopen(out, "wb");
    if (fo) {
      std::fwrite(ranks.data(), sizeof(double), ranks.size(), fo);
      std::fclose(fo);
    }
  }
}

int main(int argc, char** argv) { return ligra_main(argc, argv); }
