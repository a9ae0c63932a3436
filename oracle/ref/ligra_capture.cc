// TEST INFRASTRUCTURE ONLY -- never linked into the product library.
//
// Runs one of the reference's Ligra apps *unmodified* (its own parallel_main,
// readGraph, preprocessGraph and Compute) and writes the per-vertex arrays
// that Compute frees at its end to binary files, so that golden vectors come
// from the reference code itself rather than from a restatement.
//
// How: the app's translation unit is included below with two macros active:
//   * `Compute` is renamed so the wrapper defined here is the one that
//     parallel_main (ligra.h:559-654) calls;
//   * `free(p)` inside the app source is captured (not freed) while the
//     wrapper runs, so the arrays the app frees last stay readable.
// The arrays captured are the app's final frees, in source order:
//   BFS.C:116        free(Parents)                       -> parents  uint32[n]
//   PageRank.C:136   free(p_curr); free(p_next)          -> ranks    f64[n]
//   Radii.C:156      free(radii)                         -> radii    int32[n]
//   BC-manyiter-v2.C:235-237 free(NumPaths); free(Visited); free(Dependencies)
//                                                        -> deps     f64[n]
// Build: oracle/Makefile compiles this file once per (variant, app) with
// -DRG_APP_SOURCE="\"<App>.C\"" and the reference's apps/ and ligra/ dirs on
// the include path.  Outputs go to oracle/_ref/ only.
//
// Usage: <binary> <ligra args...> ; env RG_DUMP=<prefix> selects the output
// prefix (files <prefix>.arr<i>.bin for the i-th captured free counted from
// the end, and <prefix>.ids.bin for new_ids).  The reference's own main
// (parallel_main, ligra.h:559) runs unchanged; the dump happens in an atexit
// handler, after main has returned.

// System headers first, so the `free` macro below cannot touch their
// declarations.
#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fcntl.h>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits.h>
#include <limits>
#include <malloc.h>
#include <math.h>
#include <omp.h>
#include <parallel/algorithm>
#include <random>
#include <sstream>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <sys/types.h>
#include <tuple>
#include <type_traits>
#include <unistd.h>
#include <vector>

#include "ligra.h"

static std::vector<void*> g_captured;
static bool g_capture = false;
static long g_n = 0;
static std::vector<uintE> g_ids;

static void rg_capture_free(void* p) {
  if (g_capture) g_captured.push_back(p);
  else std::free(p);
}

#define free(p) rg_capture_free(p)
#define Compute rg_ref_compute
#include RG_APP_SOURCE
#undef Compute
#undef free

template <class vertex>
void Compute(graph<vertex>& G, commandLine P, pvector<uintE>& new_ids) {
  g_captured.clear();
  g_capture = true;
  rg_ref_compute(G, P, new_ids);
  g_capture = false;
  g_n = G.n;
  g_ids.assign(new_ids.begin(), new_ids.end());
}

static void rg_dump_at_exit() {
  const char* prefix = std::getenv("RG_DUMP");
  if (!prefix) return;
  // The caller states the element size of each captured slot, counted from
  // the last free backwards, via RG_DUMP_SIZES (e.g. "8,8").
  std::vector<size_t> sizes;
  const char* sz = std::getenv("RG_DUMP_SIZES");
  if (sz) {
    std::stringstream ss(sz);
    std::string tok;
    while (std::getline(ss, tok, ',')) sizes.push_back(std::stoul(tok));
  }
  size_t k = g_captured.size();
  for (size_t i = 0; i < sizes.size() && i < k; ++i) {
    void* p = g_captured[k - 1 - i];
    std::string fn = std::string(prefix) + ".arr" + std::to_string(i) + ".bin";
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) return;
    std::fwrite(p, sizes[i], (size_t)g_n, f);
    std::fclose(f);
  }
  std::string fn = std::string(prefix) + ".ids.bin";
  FILE* f = std::fopen(fn.c_str(), "wb");
  if (!f) return;
  std::fwrite(g_ids.data(), sizeof(uintE), g_ids.size(), f);
  std::fclose(f);
}

static int g_registered = std::atexit(rg_dump_at_exit);
