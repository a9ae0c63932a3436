// Radii with Ligra's interface (push-new-orig Ligra/apps/Radii.C:98-157):
// 64 sources hashInt(i) % n, bit-parallel frontier, radii[v] = last round in
// which v's visited word grew.  Usage: Radii [-s] [-b] [-rounds R]
// [-preprocess 0|1] [-variant push|hubdup|radar] [-out file] <graph>
#include <cstdio>

#include "ligra_rg.h"

using namespace ligra_rg;

template <class vertex>
void ligra_rg::Compute(graph<vertex>& GA, commandLine P, pvector<uintE>& new_ids) {
  Timer tm;
  tm.Start();
  bool preprocessed = (new_ids[0] != new_ids[1]);
  long n = GA.n;
  array<intE> radii(GA.ctx, n);
  array<long> Visited(GA.ctx, n), NextVisited(GA.ctx, n);
  radii.fill(-1);
  Visited.fill(0);
  NextVisited.fill(0);
  long sampleSize = std::min(n, 64L);
  std::vector<uintE> starts(sampleSize);
  for (long i = 0; i < sampleSize; ++i) {
    uintE v = (uintE)(rg_hash_int((uint64_t)i) % (uint64_t)n);
    if (preprocessed) v = new_ids[v];
    radii.set(v, 0);
    starts[i] = v;
    NextVisited.set(v, (long)1 << i);
  }
  vertexSubset Frontier(GA, n, sampleSize, starts.data());
  intE round = 0;
  while (!Frontier.isEmpty()) {
    round++;
    vertexMap(GA, Frontier, Radii_Vertex_F{Visited, NextVisited});
    vertexSubset output = edgeMap(GA, Frontier, nullptr, Radii_F{Visited, NextVisited, radii, round}, -1,
                                  dense_forward);
    Frontier.del();
    Frontier = output;
  }
  Frontier.del();
  rg_ctx_sync(GA.ctx);
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Iters until convergence = " << round << std::endl;
  if (char* out = P.getOptionValue("-out")) {
    std::vector<intE> r = radii.get();
    FILE* fo = std::fopen(out, "wb");
    if (fo) {
      std::fwrite(r.data(), sizeof(intE), r.size(), fo);
      std::fclose(fo);
    }
  }
  radii.del();
  Visited.del();
  NextVisited.del();
}

int main(int argc, char** argv) { return ligra_main(argc, argv); }
