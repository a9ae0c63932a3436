// Shared host-side helpers of the radar_graph library: error reporting and
// a small std::thread parallel-for used by the host graph builders.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "radar_graph.h"

#define RG_TRY_HOST(call)       \
  do {                          \
    int _rc = (call);           \
    if (_rc != RG_OK) return _rc; \
  } while (0)

namespace rg {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

// Number of host worker threads: RG_HOST_THREADS if set, else the hardware
// concurrency capped at 16.
int host_threads();

// Calls f(begin, end) on [lo, hi) split into contiguous chunks, one per
// worker thread.
template <class F>
void parallel_chunks(int64_t lo, int64_t hi, F f) {
  int64_t n = hi - lo;
  if (n <= 0) return;
  int t = host_threads();
  if (n < 4096 || t <= 1) {
    f(lo, hi);
    return;
  }
  int64_t nt = std::min<int64_t>(t, n);
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int64_t i = 0; i < nt; ++i) {
    int64_t b = lo + n * i / nt, e = lo + n * (i + 1) / nt;
    pool.emplace_back([=, &f] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

// Calls f(i) for every i in [lo, hi) on the worker threads, in interleaved
// blocks of `grain` (dynamic-ish balance for skewed per-item cost).
template <class F>
void parallel_for(int64_t lo, int64_t hi, F f, int64_t grain = 1024) {
  int64_t n = hi - lo;
  if (n <= 0) return;
  int t = host_threads();
  if (n <= grain || t <= 1) {
    for (int64_t i = lo; i < hi; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  int64_t nblocks = (n + grain - 1) / grain;
  int64_t nt = std::min<int64_t>(t, nblocks);
  for (int64_t w = 0; w < nt; ++w) {
    pool.emplace_back([=, &f] {
      for (int64_t b = w; b < nblocks; b += nt) {
        int64_t s = lo + b * grain, e = std::min(hi, s + grain);
        for (int64_t i = s; i < e; ++i) f(i);
      }
    });
  }
  for (auto& th : pool) th.join();
}

}  // namespace rg

// Host graph: CSR (+CSC for directed graphs) in host memory.
struct rg_hgraph {
  int64_t n = 0, m = 0;
  bool symmetric = false;
  std::vector<uint64_t> off, in_off;
  std::vector<uint32_t> nbr, in_nbr;
};
