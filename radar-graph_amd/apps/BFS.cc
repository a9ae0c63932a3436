// BFS with Ligra's interface (push-new-orig Ligra/apps/BFS.C:73-117) on the
// radar_graph engine.  Usage: BFS [-s] [-b] [-r src] [-rounds R]
// [-preprocess 0|1] [-variant push|hubdup|radar] [-k hubs] [-out file] <graph>
#include <cstdio>

#include "ligra_rg.h"

using namespace ligra_rg;

template <class vertex>
void ligra_rg::Compute(graph<vertex>& GA, commandLine P, pvector<uintE>& new_ids) {
  Timer tm;
  tm.Start();
  bool preprocessed = (new_ids[0] != new_ids[1]);
  long start = P.getOptionLongValue("-r", 0);
  if (preprocessed) start = new_ids[start];
  long n = GA.n;
  BFS_F f{array<uintE>(GA.ctx, n)};
  f.Parents.fill(UINT_E_MAX);
  f.Parents.set(start, (uintE)start);
  vertexSubset Frontier(GA, n, (uintE)start);
  int iter = 0;
  while (!Frontier.isEmpty()) {
    vertexSubset output = edgeMap(GA, Frontier, nullptr, f, -1, dense_forward);
    Frontier.del();
    Frontier = output;
    ++iter;
  }
  Frontier.del();
  rg_ctx_sync(GA.ctx);
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Num. Iters until convergence = " << iter << std::endl;
  std::vector<uintE> parents = f.Parents.get();
  long numNodes = 0;
  for (long v = 0; v < n; ++v) numNodes += parents[v] != UINT_E_MAX;
  std::cout << "[OUTPUT] No. of nodes in BFS-Tree = " << numNodes << std::endl;
  if (char* out = P.getOptionValue("-out")) {
    FILE* fo = std::fopen(out, "wb");
    if (fo) {
      std::fwrite(parents.data(), sizeof(uintE), parents.size(), fo);
      std::fclose(fo);
    }
  }
  f.Parents.del();
}

int main(int argc, char** argv) { return ligra_main(argc, argv); }
