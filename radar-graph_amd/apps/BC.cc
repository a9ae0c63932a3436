// Betweenness centrality with Ligra's interface (push-new-orig
// Ligra/apps/BC-manyiter-v2.C:124-241): -maxiters sources (default 16) from
// mt19937(27491095), forward path counting level by level, then backward
// dependency accumulation on the transposed graph; Dependencies carry over
// between sources and the transpose is not undone, as in the reference.
// Usage: BC [-s] [-b] [-maxiters K] [-rounds R] [-preprocess 0|1]
// [-variant push|hubdup|radar] [-out file] <graph>
#include <cstdio>

#include "ligra_rg.h"

using namespace ligra_rg;

template <class vertex>
void ligra_rg::Compute(graph<vertex>& GA, commandLine P, pvector<uintE>& new_ids) {
  Timer tm;
  tm.Start();
  long outputIters = 0;
  int maxIters = P.getOptionIntValue("-maxiters", 16);
  long n = GA.n;
  array<double> NumPaths(GA.ctx, n), Dependencies(GA.ctx, n);
  array<uint8_t> Visited(GA.ctx, n);
  Dependencies.fill(0.0);
  bool preprocessed = (new_ids[0] != new_ids[1]);
  std::vector<uint32_t> sources(maxIters);
  if (maxIters > 0) check(rg_pick_sources_bc(GA.host, maxIters, sources.data()), "BC sources");
  for (int iteration = 0; iteration < maxIters; ++iteration) {
    long start = sources[iteration];
    if (preprocessed) start = new_ids[start];
    NumPaths.fill(0.0);
    NumPaths.set(start, 1.0);
    Visited.fill(0);
    Visited.set(start, 1);
    vertexSubset Frontier(GA, n, (uintE)start);
    std::vector<vertexSubset> Levels;
    Levels.push_back(Frontier);
    long round = 0;
    while (!Frontier.isEmpty()) {
      round++;
      vertexSubset output = edgeMap(GA, Frontier, nullptr, BC_F{NumPaths, Visited}, -1, dense_forward);
      vertexMap(GA, output, BC_Vertex_F{Visited});
      Levels.push_back(output);
      Frontier = output;
    }
    outputIters += round;
    vertexSubset everyone = all_vertices(GA);
    vertexMap(GA, everyone, Invert_F{NumPaths});
    Levels[round].del();
    Visited.fill(0);
    Frontier = Levels[round - 1];
    vertexMap(GA, Frontier, BC_Back_Vertex_F{Visited, Dependencies, NumPaths});
    GA.transpose();
    for (long r = round - 2; r >= 0; r--) {
      vertexSubset none = edgeMap(GA, Frontier, nullptr, BC_Back_F{Dependencies, Visited}, -1,
                                  no_output | dense_forward);
      none.del();
      Frontier.del();
      Frontier = Levels[r];
      vertexMap(GA, Frontier, BC_Back_Vertex_F{Visited, Dependencies, NumPaths});
    }
    Frontier.del();
    vertexMap(GA, everyone, BC_Final_F{Dependencies, NumPaths});
    everyone.del();
  }
  rg_ctx_sync(GA.ctx);
  tm.Stop();
  tm.PrintTime("Run Time(sec) ", tm.Seconds());
  std::cout << "[OUTPUT] Combined Iters (in front traversal) = " << outputIters << std::endl;
  if (char* out = P.getOptionValue("-out")) {
    std::vector<double> d = Dependencies.get();
    FILE* fo = std::fopen(out, "wb");
    if (fo) {
      std::fwrite(d.data(), sizeof(double), d.size(), fo);
      std::fclose(fo);
    }
  }
  NumPaths.del();
  Dependencies.del();
  Visited.del();
}

int main(int argc, char** argv) { return ligra_main(argc, argv); }
