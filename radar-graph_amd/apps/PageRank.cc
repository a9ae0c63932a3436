// PageRank with Ligra's interface (push-new-orig Ligra/apps/PageRank.C:92-141).
// The iteration loop (edgeMap PR_F, vertexMap PR_Vertex_F, the L1 norm and
// the swap) runs as one fused device loop (rg_pagerank) rather than one
// host round trip per phase.  Usage: PageRank [-s] [-b] [-maxiters K]
// [-rounds R] [-preprocess 0|1] [-variant push|hubdup|radar|blocked|pull]
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
  check(rg_pagerank(GA.g, maxIters, &engine_opts(), ranks.data(), &iters, &l1, &ms), "PageRank");
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
