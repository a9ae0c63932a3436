// ligra_rg.h -- Ligra's host API on the radar_graph device engine.
//
// An app written against the reference's Ligra (push-new-orig
// Ligra/ligra/ligra.h) keeps its shape: commandLine, graph<vertex>,
// vertexSubset, edgeMap(GA, vs, nextBitmap, F(...), threshold, flags),
// vertexMap(vs, F(...)) and Compute(graph&, commandLine, pvector<uintE>&)
// driven by parallel_main's read / preprocess / rounds loop (ligra.h:559-654).
// Functors are the reference's (BFS_F, PR_F, Radii_F, BC_F, BC_Back_F and
// the vertex functors) as small structs naming a device operator and its
// state arrays; the work runs in the HIP kernels behind include/radar_graph.h.
// Errors print the library's message and abort, as the reference aborts.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "radar_graph.h"

namespace ligra_rg {

typedef uint32_t uintE;
typedef int32_t intE;
typedef int64_t intT;
typedef uint32_t flags;
const flags no_output = RG_NO_OUTPUT;      // ligra.h:52
const flags dense_forward = RG_DENSE_FORWARD; // ligra.h:55
const uintE UINT_E_MAX = 0xffffffffu;

inline void check(int rc, const char* what) {
  if (rc != RG_OK) {
    std::cout << what << ": " << rg_last_error() << std::endl;
    std::abort();
  }
}

// parseCommandLine.h:33-114: arguments are counted from the end, options
// are matched anywhere before the last argument.
class commandLine {
 public:
  int argc;
  char** argv;
  std::string comLine;
  commandLine(int c, char** v, std::string cl = "bad arguments") : argc(c), argv(v), comLine(cl) {}
  void badArgument() {
    std::cout << "usage: " << argv[0] << " " << comLine << std::endl;
    std::exit(0);
  }
  char* getArgument(int i) {
    if (argc < 2 + i) badArgument();
    return argv[argc - 1 - i];
  }
  bool getOption(const std::string& o) {
    for (int i = 1; i < argc; i++)
      if (o == argv[i]) return true;
    return false;
  }
  char* getOptionValue(const std::string& o) {
    for (int i = 1; i < argc - 1; i++)
      if (o == argv[i]) return argv[i + 1];
    return nullptr;
  }
  std::string getOptionValue(const std::string& o, const std::string& d) {
    char* v = getOptionValue(o);
    return v ? std::string(v) : d;
  }
  long getOptionLongValue(const std::string& o, long d) {
    char* v = getOptionValue(o);
    return v ? std::atol(v) : d;
  }
  int getOptionIntValue(const std::string& o, int d) { return (int)getOptionLongValue(o, d); }
  double getOptionDoubleValue(const std::string& o, double d) {
    char* v = getOptionValue(o);
    return v ? std::atof(v) : d;
  }
};

// timer.h:52-53 output format.
class Timer {
  std::chrono::steady_clock::time_point a_, b_;

 public:
  void Start() { a_ = std::chrono::steady_clock::now(); }
  void Stop() { b_ = std::chrono::steady_clock::now(); }
  double Seconds() const { return std::chrono::duration<double>(b_ - a_).count(); }
  void PrintTime(const std::string& s, double seconds) { std::printf("%-21s%3lf\n", (s + ":").c_str(), seconds); }
};

template <class T>
using pvector = std::vector<T>;

struct symmetricVertex {};
struct asymmetricVertex {};

// graph<vertex> (graph.h:57-87) on the device.
template <class vertex>
struct graph {
  rg_ctx* ctx = nullptr;
  rg_graph* g = nullptr;
  rg_hgraph* host = nullptr;   // the graph as read (original ids), for host-side pickers
  long n = 0, m = 0;
  bool transposed = false;
  void transpose() {
    check(rg_graph_transpose(g), "transpose");
    transposed = !transposed;
  }
  void del() {
    if (g) rg_graph_free(g);
    g = nullptr;
  }
};

// A device state array (the apps' newA arrays).
template <class T>
struct array {
  rg_array* a = nullptr;
  long n = 0;
  array() {}
  array(rg_ctx* ctx, long count) : n(count) { check(rg_array_alloc(ctx, count, sizeof(T), &a), "newA"); }
  void fill(T v) {
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    check(rg_array_fill(a, bits), "fill");
  }
  void set(long i, T v) { check(rg_array_write(a, &v, i, 1), "write"); }
  // swap(p_curr, p_next) of two arrays (PageRank.C:132).
  friend void swap(array& x, array& y) {
    std::swap(x.a, y.a);
    std::swap(x.n, y.n);
  }
  std::vector<T> get() const {
    std::vector<T> h(n);
    check(rg_array_read(a, h.data(), 0, n), "read");
    return h;
  }
  void del() {
    if (a) rg_array_free(a);
    a = nullptr;
  }
};

// vertexSubset (vertexSubset.h:113-229).
struct vertexSubset {
  rg_subset* s = nullptr;
  long n = 0;
  vertexSubset() {}
  template <class V>
  vertexSubset(graph<V>& G, long _n) : n(_n) {
    check(rg_subset_create(G.g, &s), "vertexSubset");
  }
  template <class V>
  vertexSubset(graph<V>& G, long _n, uintE v) : n(_n) {
    check(rg_subset_create(G.g, &s), "vertexSubset");
    check(rg_subset_set(s, &v, 1), "vertexSubset");
  }
  template <class V>
  vertexSubset(graph<V>& G, long _n, long m, const uintE* ids) : n(_n) {
    check(rg_subset_create(G.g, &s), "vertexSubset");
    check(rg_subset_set(s, ids, m), "vertexSubset");
  }
  long numNonzeros() const {
    int64_t k = 0;
    check(rg_subset_size(s, &k), "size");
    return (long)k;
  }
  long size() const { return numNonzeros(); }
  long numRows() const { return n; }
  bool isEmpty() const { return numNonzeros() == 0; }
  void del() {
    if (s) rg_subset_free(s);
    s = nullptr;
  }
};

template <class V>
vertexSubset all_vertices(graph<V>& G) {
  vertexSubset v(G, G.n);
  check(rg_subset_set_all(v.s), "vertexSubset");
  return v;
}

// The reference's edge functors as device operators.
struct BFS_F {                 // BFS.C:28-49
  array<uintE> Parents;
  static constexpr int op = RG_OP_BFS;
  std::vector<rg_array*> state() const { return {Parents.a}; }
  long round() const { return 0; }
};
struct PR_F {                  // PageRank.C:30-43 (contrib[s] = p_curr[s]/outdeg(s))
  array<double> contrib, p_next;
  static constexpr int op = RG_OP_PAGERANK;
  std::vector<rg_array*> state() const { return {contrib.a, p_next.a}; }
  long round() const { return 0; }
};
struct Radii_F {               // Radii.C:35-64
  array<long> Visited, NextVisited;
  array<intE> radii;
  intE r;
  static constexpr int op = RG_OP_RADII;
  std::vector<rg_array*> state() const { return {Visited.a, NextVisited.a, radii.a}; }
  long round() const { return r; }
};
struct BC_F {                  // BC-manyiter-v2.C:32-56
  array<double> NumPaths;
  array<uint8_t> Visited;
  static constexpr int op = RG_OP_BC_FWD;
  std::vector<rg_array*> state() const { return {NumPaths.a, Visited.a}; }
  long round() const { return 0; }
};
struct BC_Back_F {             // BC-manyiter-v2.C:58-79
  array<double> Dependencies;
  array<uint8_t> Visited;
  static constexpr int op = RG_OP_BC_BACK;
  std::vector<rg_array*> state() const { return {Dependencies.a, Visited.a}; }
  long round() const { return 0; }
};

// Engine options shared by every edgeMap of the app (-variant, -k).
inline rg_opts& engine_opts() {
  static rg_opts o = [] {
    rg_opts x;
    rg_opts_default(&x);
    return x;
  }();
  return o;
}

// edgeMap (ligra.h:336-340).  The bool* nextBitmap argument of the
// reference is accepted and ignored: the next frontier lives on the device.
template <class V, class F>
vertexSubset edgeMap(graph<V>& GA, vertexSubset& vs, bool* /*nextBitmap*/, F f, intT threshold = -1,
                     const flags& fl = 0) {
  vertexSubset out(GA, GA.n);
  std::vector<rg_array*> st = f.state();
  check(rg_edgemap(GA.g, F::op, vs.s, st.data(), (int)st.size(), f.round(), threshold, (int)fl, &engine_opts(),
                   out.s),
        "edgeMap");
  return out;
}

// Vertex functors (vertexMap, ligra.h:464-479).
struct PR_Vertex_F {           // PageRank.C:46-58
  array<double> p_next;
  double damping, added;
  static constexpr int vop = RG_VOP_PR_UPDATE;
  std::vector<rg_array*> state() const { return {p_next.a}; }
};
struct Radii_Vertex_F {        // Radii.C:67-75
  array<long> Visited, NextVisited;
  static constexpr int vop = RG_VOP_RADII_SYNC;
  std::vector<rg_array*> state() const { return {Visited.a, NextVisited.a}; }
};
struct BC_Vertex_F {           // BC-manyiter-v2.C:82-89
  array<uint8_t> Visited;
  static constexpr int vop = RG_VOP_MARK;
  std::vector<rg_array*> state() const { return {Visited.a}; }
};
struct BC_Back_Vertex_F {      // BC-manyiter-v2.C:93-101
  array<uint8_t> Visited;
  array<double> Dependencies, inverseNumPaths;
  static constexpr int vop = RG_VOP_BC_BACK;
  std::vector<rg_array*> state() const { return {Visited.a, Dependencies.a, inverseNumPaths.a}; }
};

struct Invert_F {              // inverseNumPaths[i] = 1/NumPaths[i] (BC-manyiter-v2.C:206-208)
  array<double> x;
  static constexpr int vop = RG_VOP_INVERT;
  std::vector<rg_array*> state() const { return {x.a}; }
};
struct BC_Final_F {            // Dependencies = (Dep - inv)/inv (BC-manyiter-v2.C:230-232)
  array<double> Dependencies, inverseNumPaths;
  static constexpr int vop = RG_VOP_BC_FINAL;
  std::vector<rg_array*> state() const { return {Dependencies.a, inverseNumPaths.a}; }
};

struct PR_Vertex_Contrib_F {   // contrib[i] = p[i] / outdeg(i): PR_F's quotient (PageRank.C:36-40)
  array<double> p, contrib;
  static constexpr int vop = RG_VOP_PR_CONTRIB;
  std::vector<rg_array*> state() const { return {p.a, contrib.a}; }
};
struct Abs_Diff_F {            // x[i] = fabs(x[i] - y[i]) (PageRank.C:125-127)
  array<double> x, y;
  static constexpr int vop = RG_VOP_ABS_DIFF;
  std::vector<rg_array*> state() const { return {x.a, y.a}; }
};
struct Abs_F {                 // x[i] = fabs(y[i]) (PageRankDelta.C:136-137)
  array<double> x, y;
  static constexpr int vop = RG_VOP_ABS;
  std::vector<rg_array*> state() const { return {x.a, y.a}; }
};

template <class V, class F>
void vertexMap(graph<V>& GA, vertexSubset& vs, F f, double a = 0.0, double b = 0.0) {
  std::vector<rg_array*> st = f.state();
  check(rg_vertexmap(GA.g, F::vop, vs.s, st.data(), (int)st.size(), a, b), "vertexMap");
}

// Filter functors (vertexFilter, ligra.h:500-509).
struct PR_Vertex_F_FirstRound {   // PageRankDelta.C:49-63
  array<double> p, Delta, nghSum;
  double damping, one_over_n, epsilon2;
  static constexpr int fop = RG_FOP_PRD_FIRST;
  std::vector<rg_array*> state() const { return {p.a, Delta.a, nghSum.a}; }
  std::vector<double> params() const { return {damping, one_over_n, epsilon2}; }
};
struct PR_Delta_Vertex_F {        // PR_Vertex_F of PageRankDelta.C:65-77
  array<double> p, Delta, nghSum;
  double damping, epsilon2;
  static constexpr int fop = RG_FOP_PRD;
  std::vector<rg_array*> state() const { return {p.a, Delta.a, nghSum.a}; }
  std::vector<double> params() const { return {damping, epsilon2}; }
};
struct Unvisited_F {              // keeps Parents[i] == UINT_E_MAX (BFS.C:47)
  array<uintE> Parents;
  static constexpr int fop = RG_FOP_U32_EQ;
  std::vector<rg_array*> state() const { return {Parents.a}; }
  std::vector<double> params() const { return {(double)UINT_E_MAX}; }
};

template <class V, class F>
vertexSubset vertexFilter(graph<V>& GA, vertexSubset& vs, F f) {
  vertexSubset out(GA, GA.n);
  std::vector<rg_array*> st = f.state();
  std::vector<double> pr = f.params();
  check(rg_vertexfilter(GA.g, F::fop, vs.s, st.data(), (int)st.size(), pr.data(), (int)pr.size(), out.s),
        "vertexFilter");
  return out;
}

// sequence::plusReduce over a device array of doubles.
inline double plusReduce(const array<double>& x) {
  double s = 0.0;
  check(rg_array_sum(x.a, &s), "plusReduce");
  return s;
}

template <class vertex>
void Compute(graph<vertex>& G, commandLine P, pvector<uintE>& new_ids);

// parallel_main (ligra.h:559-654): read the graph (-s symmetric, -b Ligra
// binary; otherwise AdjacencyGraph text; files ending in .el/.wel/.sg/.gr/
// .graph/.mtx use the GAP readers), optionally degree-sort it on the GPU
// (-preprocess 0 out-degree,
// 1 in-degree; the DS/RADAR default is 1), then run Compute once plus
// -rounds more times, transposing a directed graph back between rounds.
// Extensions: -variant push|hubdup|radar|blocked|pull, -k <hub slots>,
// -device <i>.
inline int ligra_main(int argc, char** argv, int default_preprocess = -1) {
  commandLine P(argc, argv, " [-s] <inFile>");
  char* iFile = P.getArgument(0);
  bool symmetric = P.getOptionValue("-s") != nullptr;
  bool binary = P.getOptionValue("-b") != nullptr;
  int preprocess = P.getOptionIntValue("-preprocess", default_preprocess);
  long rounds = P.getOptionLongValue("-rounds", 3);
  std::string variant = P.getOptionValue("-variant", "push");
  rg_opts& eo = engine_opts();
  eo.variant = variant == "hubdup" ? RG_HUBDUP
               : variant == "radar" ? RG_RADAR
               : variant == "blocked" ? RG_BLOCKED
               : variant == "pull" ? RG_PULL : RG_PUSH;
  eo.hubs = P.getOptionLongValue("-k", 0);
  std::string f(iFile);
  int fmt = binary ? RG_FMT_ADJ_BINARY : rg_format_of(f.c_str());
  Timer t;
  t.Start();
  graph<symmetricVertex> G;
  check(rg_hgraph_read(iFile, fmt, symmetric ? 1 : 0, &G.host), "read");
  int64_t n = 0, m = 0;
  int sym = 0;
  rg_hgraph_info(G.host, &n, &m, &sym);
  t.Stop();
  t.PrintTime("Graph reading time(s)", t.Seconds());
  std::cout << "Read " << (sym ? "undirected" : "directed") << " graph. Num Nodes = " << n
            << " and Num Edges = " << m << "\n";
  check(rg_ctx_create(P.getOptionIntValue("-device", 0), &G.ctx), "device");
  check(rg_graph_upload(G.ctx, G.host, &G.g), "upload");
  G.n = n;
  G.m = m;
  pvector<uintE> new_ids(n, 0);
  if (preprocess == 0 || preprocess == 1) {
    t.Start();
    rg_graph* sorted = nullptr;
    check(rg_graph_degree_sort(G.g, preprocess == 0, 0, &sorted, new_ids.data()), "DegSort");
    rg_ctx_sync(G.ctx);
    t.Stop();
    t.PrintTime("DegSort Time", t.Seconds());
    rg_graph_free(G.g);
    G.g = sorted;
  }
  for (long r = 0; r <= rounds; ++r) {
    Compute(G, P, new_ids);
    if (G.transposed) G.transpose();
  }
  G.del();
  rg_hgraph_free(G.host);
  rg_ctx_destroy(G.ctx);
  return 0;
}

}  // namespace ligra_rg
