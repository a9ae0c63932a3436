// TEST INFRASTRUCTURE ONLY -- never linked into the product library.
//
// Drives the reference's GAP graph code (generator.h, builder.h,
// benchmark.h) unmodified and writes what it produces, so the product's
// generator / CSR builder / source picker can be checked against it:
//
//   gap_graph_dump el  <scale> <degree> <uniform 0|1> <out>
//       raw edge list of Generator::GenerateEL (generator.h:64-124),
//       int32 pairs (u, v), before any CSR building;
//   gap_graph_dump csr <scale> <degree> <uniform 0|1> <out>
//       Builder::MakeGraph (builder.h:226-244) with the symmetrize that -g/-u
//       imply (command_line.h:77-78): int64 n, int64 m, int64 offsets[n+1],
//       int32 neighbours[m] of the squished (sorted, deduplicated) CSR;
//   gap_graph_dump relabel <scale> <degree> <uniform 0|1> <out>
//       the same CSR after Builder::RelabelByDegree (builder.h:246-282);
//   gap_graph_dump sources <scale> <degree> <uniform 0|1> <count> <out>
//       the first <count> picks of SourcePicker (benchmark.h:47-68) on the
//       built graph, int32 each.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "benchmark.h"
#include "builder.h"
#include "command_line.h"
#include "generator.h"
#include "graph.h"

static void write_csr(const Graph& g, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) std::exit(3);
  int64_t n = g.num_nodes();
  int64_t m = g.num_edges_directed();
  std::fwrite(&n, 8, 1, f);
  std::fwrite(&m, 8, 1, f);
  pvector<SGOffset> off = g.VertexOffsets(false);
  std::fwrite(off.data(), sizeof(SGOffset), n + 1, f);
  std::fwrite(g.out_neigh(0).begin(), sizeof(NodeID), m, f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  std::string mode = argv[1];
  int scale = std::atoi(argv[2]);
  int degree = std::atoi(argv[3]);
  bool uniform = std::atoi(argv[4]) != 0;
  if (mode == "el") {
    Generator<NodeID> gen(scale, degree);
    pvector<EdgePair<NodeID, NodeID>> el = gen.GenerateEL(uniform);
    FILE* f = std::fopen(argv[5], "wb");
    if (!f) return 3;
    for (auto& e : el) {
      int32_t uv[2] = {e.u, e.v};
      std::fwrite(uv, 4, 2, f);
    }
    std::fclose(f);
    return 0;
  }
  std::string sarg = std::to_string(scale), karg = std::to_string(degree);
  std::vector<std::string> args = {"gap", uniform ? "-u" : "-g", sarg, "-k", karg};
  std::vector<char*> cargv;
  for (auto& a : args) cargv.push_back(&a[0]);
  CLBase cli((int)cargv.size(), cargv.data());
  if (!cli.ParseArgs()) return 2;
  Builder b(cli);
  Graph g = b.MakeGraph();
  if (mode == "csr") {
    write_csr(g, argv[5]);
  } else if (mode == "relabel") {
    Graph r = Builder::RelabelByDegree(g);
    write_csr(r, argv[5]);
  } else if (mode == "sources") {
    if (argc < 7) return 2;
    int count = std::atoi(argv[5]);
    SourcePicker<Graph> sp(g);
    FILE* f = std::fopen(argv[6], "wb");
    if (!f) return 3;
    for (int i = 0; i < count; ++i) {
      int32_t s = sp.PickNext();
      std::fwrite(&s, 4, 1, f);
    }
    std::fclose(f);
  } else {
    return 2;
  }
  return 0;
}
