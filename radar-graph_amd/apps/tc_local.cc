// Triangle counting with GAP's command line (push-new-orig GAP/src/
// tc_local.cc:151-163, command_line.h:28-139): -g <scale> R-MAT or
// -u <scale> uniform (both symmetrised), -k <degree> (default 16),
// -f <file> (.el / .wel / .sg / .gr / .graph / .mtx / Ligra .adj), -s symmetrize, -n <trials>, -v verify,
// -a analysis (prints the count).  -r relabels by degree first (the
// degree-sorting and RADAR trees, tc_local.cc:116 / 158); the relabel is
// timed separately from the count.
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ligra_rg.h"

using ligra_rg::check;

static bool verify(const rg_hgraph* hg, uint64_t total) {
  // TCVerifier (tc_local.cc:129-148): sum over arcs (u, v) of
  // |N(u) & N(v)|, each triangle counted six times.
  int64_t n = 0, m = 0;
  int sym = 0;
  rg_hgraph_info(hg, &n, &m, &sym);
  const uint64_t* off = rg_hgraph_offsets(hg, 0);
  const uint32_t* nb = rg_hgraph_neighbors(hg, 0);
  uint64_t t = 0;
  std::vector<uint32_t> tmp;
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t a = off[u]; a < off[u + 1]; ++a) {
      uint32_t v = nb[a];
      tmp.clear();
      std::set_intersection(nb + off[u], nb + off[u + 1], nb + off[v], nb + off[v + 1], std::back_inserter(tmp));
      t += tmp.size();
    }
  return t / 6 == total;
}

int main(int argc, char** argv) {
  int scale = -1, degree = 16, trials = 1;
  bool uniform = false, symmetrize = false, do_verify = false, analysis = false, relabel = false;
  std::string file;
  int c;
  while ((c = getopt(argc, argv, "f:g:u:k:sn:var")) != -1) {
    switch (c) {
      case 'f': file = optarg; break;
      case 'g': scale = std::atoi(optarg); break;
      case 'u': scale = std::atoi(optarg); uniform = true; break;
      case 'k': degree = std::atoi(optarg); break;
      case 's': symmetrize = true; break;
      case 'n': trials = std::atoi(optarg); break;
      case 'v': do_verify = true; break;
      case 'a': analysis = true; break;
      case 'r': relabel = true; break;
      default: break;
    }
  }
  if (file.empty() && scale == -1) {
    std::printf("No graph input specified. (Use -h for help)\n");
    return -1;
  }
  if (scale != -1) symmetrize = true;
  rg_hgraph* hg = nullptr;
  if (!file.empty()) {
    int fmt = rg_format_of(file.c_str());
    check(rg_hgraph_read(file.c_str(), fmt, symmetrize ? 1 : 0, &hg), "read");
  } else {
    check(rg_hgraph_generate(scale, degree, uniform ? 1 : 0, 1, &hg), "generate");
  }
  int64_t n = 0, m = 0;
  int sym = 0;
  rg_hgraph_info(hg, &n, &m, &sym);
  if (!sym) {
    std::printf("Input graph is directed but tc requires undirected\n");
    return -2;
  }
  std::printf("Graph has %lld nodes and %lld undirected edges for degree: %lld\n", (long long)n,
              (long long)(m / 2), (long long)(n ? m / n : 0));
  rg_ctx* ctx = nullptr;
  check(rg_ctx_create(0, &ctx), "device");
  rg_graph* g = nullptr;
  check(rg_graph_upload(ctx, hg, &g), "upload");
  if (relabel) {
    ligra_rg::Timer t;
    t.Start();
    rg_graph* r = nullptr;
    check(rg_graph_degree_sort(g, 1, 1, &r, nullptr), "relabel");
    rg_ctx_sync(ctx);
    t.Stop();
    std::printf("%-21s%3.5lf\n", "Relabel:", t.Seconds());
    rg_graph_free(g);
    g = r;
  }
  uint64_t total = 0;
  double total_s = 0.0;
  for (int i = 0; i < trials; ++i) {
    double ms = 0.0;
    check(rg_tc(g, nullptr, &total, nullptr, &ms), "tc");
    std::printf("Run Time (in s) = %g\n", ms / 1e3);
    std::printf("%-21s%3.5lf\n", "Trial Time:", ms / 1e3);
    total_s += ms / 1e3;
    if (analysis && i == trials - 1) std::printf("%llu triangles\n", (unsigned long long)total);
    if (do_verify) std::printf("%-21s%7s\n", "Verification:", verify(hg, total) ? "PASS" : "FAIL");
  }
  std::printf("%-21s%3.5lf\n", "Average Time:", total_s / trials);
  rg_graph_free(g);
  rg_hgraph_free(hg);
  rg_ctx_destroy(ctx);
  return 0;
}
