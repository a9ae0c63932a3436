// Host-side graph ingest for the radar_graph library: the GAP generator,
// CSR construction with the reference readers' dedup semantics, the Ligra and
// GAP file formats, and the reference's source pickers.
//
// Reference behaviour followed (paths relative to the reference's src/):
//   generator  push-new-orig/GAP/src/generator.h:52-124
//   CSR build  push-new-orig/GAP/src/builder.h:80-224 and
//              push-new-orig/Ligra/ligra/IO.h:364-475 (sort, unique, drop self-loops)
//   readers    push-new-orig/Ligra/ligra/IO.h:207-626, GAP/src/reader.h:50-304
//   writers    push-new-orig/GAP/src/writer.h:32-69
//   pickers    push-new-orig/GAP/src/benchmark.h:47-68, Ligra/ligra/utils.h:347-355
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>

#include "rg_common.h"

namespace rg {

static thread_local std::string t_last_error;

int set_error(int code, const std::string& msg) {
  t_last_error = msg;
  return code;
}

int host_threads() {
  static int t = [] {
    const char* e = std::getenv("RG_HOST_THREADS");
    if (e && std::atoi(e) > 0) return std::atoi(e);
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h));
  }();
  return t;
}

namespace {

constexpr int64_t kRandSeed = 27491095;        // GAP util.h:22
constexpr int64_t kGenBlock = int64_t(1) << 18; // GAP generator.h:152

// Exclusive prefix sum of 32-bit counts into 64-bit offsets (n+1 entries).
std::vector<uint64_t> exclusive_scan(const std::vector<uint32_t>& cnt) {
  int64_t n = (int64_t)cnt.size();
  std::vector<uint64_t> off(n + 1);
  int t = host_threads();
  int64_t nb = std::max<int64_t>(1, std::min<int64_t>(t, n / 65536 + 1));
  std::vector<uint64_t> part(nb + 1, 0);
  parallel_for(0, nb, [&](int64_t b) {
    int64_t s = n * b / nb, e = n * (b + 1) / nb;
    uint64_t acc = 0;
    for (int64_t i = s; i < e; ++i) acc += cnt[i];
    part[b + 1] = acc;
  }, 1);
  for (int64_t b = 0; b < nb; ++b) part[b + 1] += part[b];
  parallel_for(0, nb, [&](int64_t b) {
    int64_t s = n * b / nb, e = n * (b + 1) / nb;
    uint64_t acc = part[b];
    for (int64_t i = s; i < e; ++i) { off[i] = acc; acc += cnt[i]; }
  }, 1);
  off[n] = part[nb];
  return off;
}

// Buckets arcs (a[i] -> b[i]) by a into a CSR whose lists are then sorted;
// when `dedup`, duplicates and self-loops are removed afterwards.
// Arcs with a == b are kept unless dedup.
void bucket_sort_csr(int64_t n, int64_t narcs, const uint32_t* a, const uint32_t* b,
                     bool both_dirs, bool dedup, std::vector<uint64_t>& off_out,
                     std::vector<uint32_t>& nbr_out) {
  std::unique_ptr<std::atomic<uint32_t>[]> deg(new std::atomic<uint32_t>[n]);
  parallel_for(0, n, [&](int64_t i) { deg[i].store(0, std::memory_order_relaxed); }, 1 << 16);
  parallel_chunks(0, narcs, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      if (dedup && a[i] == b[i]) continue;
      deg[a[i]].fetch_add(1, std::memory_order_relaxed);
      if (both_dirs) deg[b[i]].fetch_add(1, std::memory_order_relaxed);
    }
  });
  std::vector<uint32_t> cnt(n);
  parallel_for(0, n, [&](int64_t i) { cnt[i] = deg[i].load(std::memory_order_relaxed); }, 1 << 16);
  std::vector<uint64_t> off = exclusive_scan(cnt);
  std::unique_ptr<std::atomic<uint64_t>[]> cur(new std::atomic<uint64_t>[n]);
  parallel_for(0, n, [&](int64_t i) { cur[i].store(off[i], std::memory_order_relaxed); }, 1 << 16);
  std::vector<uint32_t> nbr(off[n]);
  parallel_chunks(0, narcs, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      if (dedup && a[i] == b[i]) continue;
      nbr[cur[a[i]].fetch_add(1, std::memory_order_relaxed)] = b[i];
      if (both_dirs) nbr[cur[b[i]].fetch_add(1, std::memory_order_relaxed)] = a[i];
    }
  });
  cur.reset();
  deg.reset();
  // Sort every list; count the unique, self-loop-free survivors.
  parallel_for(0, n, [&](int64_t v) {
    uint32_t* s = nbr.data() + off[v];
    uint32_t* e = nbr.data() + off[v + 1];
    std::sort(s, e);
    if (dedup) {
      uint32_t* ne = std::unique(s, e);
      ne = std::remove(s, ne, (uint32_t)v);
      cnt[v] = (uint32_t)(ne - s);
    }
  }, 256);
  if (!dedup) {
    off_out.swap(off);
    nbr_out.swap(nbr);
    return;
  }
  std::vector<uint64_t> noff = exclusive_scan(cnt);
  std::vector<uint32_t> nn(noff[n]);
  parallel_for(0, n, [&](int64_t v) {
    std::copy(nbr.data() + off[v], nbr.data() + off[v] + cnt[v], nn.data() + noff[v]);
  }, 1024);
  off_out.swap(noff);
  nbr_out.swap(nn);
}

// In-edge CSR of a directed CSR: lists sorted ascending (duplicates kept).
void transpose_csr(int64_t n, const std::vector<uint64_t>& off, const std::vector<uint32_t>& nbr,
                   std::vector<uint64_t>& toff, std::vector<uint32_t>& tnbr) {
  int64_t m = (int64_t)nbr.size();
  std::vector<uint32_t> src(m);
  parallel_for(0, n, [&](int64_t v) {
    for (uint64_t j = off[v]; j < off[v + 1]; ++j) src[j] = (uint32_t)v;
  }, 1024);
  bucket_sort_csr(n, m, nbr.data(), src.data(), false, false, toff, tnbr);
}

void finish_graph(rg_hgraph* g) {
  g->m = (int64_t)g->nbr.size();
  if (!g->symmetric) {
    transpose_csr(g->n, g->off, g->nbr, g->in_off, g->in_nbr);
  } else {
    g->in_off.clear();
    g->in_nbr.clear();
  }
}

}  // namespace
}  // namespace rg

using rg::set_error;

extern "C" {

const char* rg_last_error(void) { return rg::t_last_error.c_str(); }

int64_t rg_generate_edge_count(int scale, int degree) {
  if (scale < 1 || scale > 31 || degree < 1) return -1;
  return (int64_t(1) << scale) * degree;
}

// GAP Generator::MakeRMatEL / MakeUniformEL (generator.h:64-114): blocks of
// 2^18 edges, each with its own mt19937 seeded kRandSeed + block index; R-MAT
// descends `scale` quadrant levels with one uniform float per level; then
// PermuteIDs (generator.h:52-62) relabels by a std::shuffle permutation.
int rg_generate_edges(int scale, int degree, int uniform, uint32_t* src, uint32_t* dst) {
  int64_t ne = rg_generate_edge_count(scale, degree);
  if (ne < 0 || !src || !dst) return set_error(RG_EINVAL, "rg_generate_edges: bad scale/degree/pointer");
  const int64_t nv = int64_t(1) << scale;
  const int64_t nblocks = (ne + rg::kGenBlock - 1) / rg::kGenBlock;
  if (uniform) {
    rg::parallel_for(0, nblocks, [&](int64_t blk) {
      std::mt19937 rng;
      std::uniform_int_distribution<int32_t> udist(0, (int32_t)(nv - 1));
      rng.seed(rg::kRandSeed + blk);
      int64_t s = blk * rg::kGenBlock, e = std::min(s + rg::kGenBlock, ne);
      for (int64_t i = s; i < e; ++i) {
        // GAP writes Edge(udist(rng), udist(rng)); g++ evaluates those two
        // constructor arguments right to left, so the first draw is v.
        int32_t v = udist(rng);
        int32_t u = udist(rng);
        src[i] = (uint32_t)u;
        dst[i] = (uint32_t)v;
      }
    }, 1);
    return RG_OK;
  }
  const float A = 0.57f, B = 0.19f, C = 0.19f;
  rg::parallel_for(0, nblocks, [&](int64_t blk) {
    std::mt19937 rng;
    std::uniform_real_distribution<float> udist(0, 1.0f);
    rng.seed(rg::kRandSeed + blk);
    int64_t s = blk * rg::kGenBlock, e = std::min(s + rg::kGenBlock, ne);
    for (int64_t i = s; i < e; ++i) {
      int32_t u = 0, v = 0;
      for (int depth = 0; depth < scale; ++depth) {
        float r = udist(rng);
        u <<= 1;
        v <<= 1;
        if (r < A + B) {
          if (r > A) v++;
        } else {
          u++;
          if (r > A + B + C) v++;
        }
      }
      src[i] = (uint32_t)u;
      dst[i] = (uint32_t)v;
    }
  }, 1);
  std::vector<int32_t> perm(nv);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937 prng(rg::kRandSeed);
  std::shuffle(perm.begin(), perm.end(), prng);
  rg::parallel_chunks(0, ne, [&](int64_t s, int64_t e) {
    for (int64_t i = s; i < e; ++i) {
      src[i] = (uint32_t)perm[src[i]];
      dst[i] = (uint32_t)perm[dst[i]];
    }
  });
  return RG_OK;
}

int rg_hgraph_from_edges(int64_t n, int64_t nedges, const uint32_t* src, const uint32_t* dst,
                         int symmetric, rg_hgraph** out) {
  if (!out || n <= 0 || n > (int64_t)UINT32_MAX || nedges < 0 || (nedges > 0 && (!src || !dst)))
    return set_error(RG_EINVAL, "rg_hgraph_from_edges: bad arguments");
  for (int64_t i = 0; i < nedges; ++i)
    if (src[i] >= (uint64_t)n || dst[i] >= (uint64_t)n)
      return set_error(RG_EINVAL, "rg_hgraph_from_edges: vertex id out of range");
  try {
    std::unique_ptr<rg_hgraph> g(new rg_hgraph);
    g->n = n;
    g->symmetric = symmetric != 0;
    rg::bucket_sort_csr(n, nedges, src, dst, g->symmetric, true, g->off, g->nbr);
    rg::finish_graph(g.get());
    *out = g.release();
    return RG_OK;
  } catch (const std::bad_alloc&) {
    return set_error(RG_ENOMEM, "rg_hgraph_from_edges: out of host memory");
  }
}

int rg_hgraph_generate(int scale, int degree, int uniform, int symmetric, rg_hgraph** out) {
  int64_t ne = rg_generate_edge_count(scale, degree);
  if (ne < 0 || !out) return set_error(RG_EINVAL, "rg_hgraph_generate: bad scale/degree");
  try {
    std::vector<uint32_t> s(ne), d(ne);
    int rc = rg_generate_edges(scale, degree, uniform, s.data(), d.data());
    if (rc) return rc;
    return rg_hgraph_from_edges(int64_t(1) << scale, ne, s.data(), d.data(), symmetric, out);
  } catch (const std::bad_alloc&) {
    return set_error(RG_ENOMEM, "rg_hgraph_generate: out of host memory");
  }
}

int rg_hgraph_from_csr(int64_t n, int64_t m, const uint64_t* offsets, const uint32_t* nbrs,
                       int symmetric, int dedup, rg_hgraph** out) {
  if (!out || n <= 0 || n > (int64_t)UINT32_MAX || m < 0 || !offsets || (m > 0 && !nbrs))
    return set_error(RG_EINVAL, "rg_hgraph_from_csr: bad arguments");
  if (offsets[0] != 0 || offsets[n] != (uint64_t)m)
    return set_error(RG_EFORMAT, "rg_hgraph_from_csr: offsets must start at 0 and end at m");
  for (int64_t v = 0; v < n; ++v)
    if (offsets[v + 1] < offsets[v]) return set_error(RG_EFORMAT, "rg_hgraph_from_csr: offsets decrease");
  for (int64_t j = 0; j < m; ++j)
    if (nbrs[j] >= (uint64_t)n) return set_error(RG_EFORMAT, "rg_hgraph_from_csr: neighbour out of range");
  try {
    std::unique_ptr<rg_hgraph> g(new rg_hgraph);
    g->n = n;
    g->symmetric = symmetric != 0;
    if (dedup) {
      std::vector<uint32_t> src(m);
      rg::parallel_for(0, n, [&](int64_t v) {
        for (uint64_t j = offsets[v]; j < offsets[v + 1]; ++j) src[j] = (uint32_t)v;
      }, 1024);
      // The reader dedups each list in place; the symmetric reader does not
      // add reverse arcs (IO.h:432-467), so neither do we here.
      rg::bucket_sort_csr(n, m, src.data(), nbrs, false, true, g->off, g->nbr);
    } else {
      g->off.assign(offsets, offsets + n + 1);
      g->nbr.assign(nbrs, nbrs + m);
    }
    rg::finish_graph(g.get());
    *out = g.release();
    return RG_OK;
  } catch (const std::bad_alloc&) {
    return set_error(RG_ENOMEM, "rg_hgraph_from_csr: out of host memory");
  }
}

int rg_hgraph_info(const rg_hgraph* g, int64_t* n, int64_t* m, int* symmetric) {
  if (!g) return set_error(RG_EINVAL, "rg_hgraph_info: null graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (symmetric) *symmetric = g->symmetric ? 1 : 0;
  return RG_OK;
}

const uint64_t* rg_hgraph_offsets(const rg_hgraph* g, int in_edges) {
  if (!g) return nullptr;
  return (in_edges && !g->symmetric) ? g->in_off.data() : g->off.data();
}

const uint32_t* rg_hgraph_neighbors(const rg_hgraph* g, int in_edges) {
  if (!g) return nullptr;
  return (in_edges && !g->symmetric) ? g->in_nbr.data() : g->nbr.data();
}

void rg_hgraph_free(rg_hgraph* g) { delete g; }

// ---- file formats -------------------------------------------------------

static bool read_file(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == buf.size();
}

// Whitespace-separated tokens of a text buffer, as Ligra's stringToWords
// splits on any whitespace (IO.h:92-125).
static void split_words(const std::string& buf, std::vector<std::pair<size_t, size_t>>& words) {
  size_t i = 0, n = buf.size();
  while (i < n) {
    while (i < n && std::isspace((unsigned char)buf[i])) ++i;
    if (i >= n) break;
    size_t s = i;
    while (i < n && !std::isspace((unsigned char)buf[i])) ++i;
    words.emplace_back(s, i - s);
  }
}

static bool parse_u64(const std::string& buf, std::pair<size_t, size_t> w, uint64_t& v) {
  v = 0;
  if (w.second == 0) return false;
  for (size_t k = 0; k < w.second; ++k) {
    char c = buf[w.first + k];
    if (c < '0' || c > '9') return false;
    v = v * 10 + (uint64_t)(c - '0');
  }
  return true;
}

static int read_adj_text(const char* path, int symmetric, rg_hgraph** out) {
  std::string buf;
  if (!read_file(path, buf)) return set_error(RG_EIO, std::string("cannot read ") + path);
  std::vector<std::pair<size_t, size_t>> w;
  split_words(buf, w);
  if (w.empty() || buf.compare(w[0].first, w[0].second, "AdjacencyGraph") != 0 ||
      w[0].second != 14)
    return set_error(RG_EFORMAT, "Bad input file");
  uint64_t n = 0, m = 0;
  if (w.size() < 3 || !parse_u64(buf, w[1], n) || !parse_u64(buf, w[2], m))
    return set_error(RG_EFORMAT, "Bad input file");
  if (w.size() - 1 != n + m + 2 || n == 0) return set_error(RG_EFORMAT, "Bad input file");
  std::vector<uint64_t> off(n + 1);
  std::vector<uint32_t> nbr(m);
  for (uint64_t i = 0; i < n; ++i) {
    if (!parse_u64(buf, w[3 + i], off[i])) return set_error(RG_EFORMAT, "Bad input file");
  }
  off[n] = m;
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t x;
    if (!parse_u64(buf, w[3 + n + i], x) || x >= n) return set_error(RG_EFORMAT, "Bad input file");
    nbr[i] = (uint32_t)x;
  }
  return rg_hgraph_from_csr((int64_t)n, (int64_t)m, off.data(), nbr.data(), symmetric, 1, out);
}

static int read_adj_binary(const char* prefix, int symmetric, rg_hgraph** out) {
  std::string cfg, adj, idx;
  std::string p(prefix);
  if (!read_file((p + ".config").c_str(), cfg) || !read_file((p + ".adj").c_str(), adj) ||
      !read_file((p + ".idx").c_str(), idx))
    return set_error(RG_EIO, "cannot read " + p + ".config/.adj/.idx");
  long long n = std::atoll(cfg.c_str());
  if (n <= 0) return set_error(RG_EFORMAT, "Bad input file");
  if ((long long)(idx.size() / sizeof(uint32_t)) != n) return set_error(RG_EFORMAT, "File size wrong");
  int64_t m = (int64_t)(adj.size() / sizeof(uint32_t));
  std::vector<uint64_t> off(n + 1);
  const uint32_t* io = reinterpret_cast<const uint32_t*>(idx.data());
  for (long long i = 0; i < n; ++i) off[i] = io[i];
  off[n] = (uint64_t)m;
  return rg_hgraph_from_csr(n, m, off.data(), reinterpret_cast<const uint32_t*>(adj.data()),
                            symmetric, 0, out);
}

static int read_el(const char* path, int symmetric, rg_hgraph** out) {
  std::string buf;
  if (!read_file(path, buf)) return set_error(RG_EIO, std::string("cannot read ") + path);
  std::vector<std::pair<size_t, size_t>> w;
  split_words(buf, w);
  if (w.size() % 2) return set_error(RG_EFORMAT, "edge list has an odd number of ids");
  std::vector<uint32_t> s(w.size() / 2), d(w.size() / 2);
  uint64_t maxid = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    uint64_t a, b;
    if (!parse_u64(buf, w[2 * i], a) || !parse_u64(buf, w[2 * i + 1], b) || a > INT32_MAX ||
        b > INT32_MAX)
      return set_error(RG_EFORMAT, "Bad input file");
    s[i] = (uint32_t)a;
    d[i] = (uint32_t)b;
    maxid = std::max(maxid, std::max(a, b));
  }
  // GAP: num_nodes = max id + 1 (builder.h:209-210).
  return rg_hgraph_from_edges((int64_t)maxid + 1, (int64_t)s.size(), s.data(), d.data(), symmetric, out);
}

static int read_sg(const char* path, rg_hgraph** out) {
  std::string buf;
  if (!read_file(path, buf)) return set_error(RG_EIO, std::string("cannot read ") + path);
  const size_t hdr = 1 + 8 + 8;
  if (buf.size() < hdr) return set_error(RG_EFORMAT, "Bad input file");
  bool directed = buf[0] != 0;
  int64_t m, n;
  std::memcpy(&m, buf.data() + 1, 8);
  std::memcpy(&n, buf.data() + 9, 8);
  if (n <= 0 || m < 0) return set_error(RG_EFORMAT, "Bad input file");
  size_t part = (size_t)(n + 1) * 8 + (size_t)m * 4;
  if (buf.size() < hdr + part * (directed ? 2 : 1)) return set_error(RG_EFORMAT, "Bad input file");
  std::vector<uint64_t> off(n + 1);
  std::memcpy(off.data(), buf.data() + hdr, (n + 1) * 8);
  const uint32_t* nb = reinterpret_cast<const uint32_t*>(buf.data() + hdr + (n + 1) * 8);
  int rc = rg_hgraph_from_csr(n, m, off.data(), nb, directed ? 0 : 1, 0, out);
  return rc;
}

// ---- GAP's other text readers (reader.h:59-217) -----------------------
// Each yields an edge list that the builder turns into a CSR with
// num_nodes = max id + 1 (builder.h:209-210); weights are read and dropped
// (the path's graphs are unweighted).

// Splits buf into lines and each line into whitespace-separated tokens.
struct TextLines {
  const std::string& buf;
  size_t pos = 0;
  explicit TextLines(const std::string& b) : buf(b) {}
  bool next(std::string& line) {
    if (pos >= buf.size()) return false;
    size_t e = buf.find('\n', pos);
    if (e == std::string::npos) e = buf.size();
    line.assign(buf, pos, e - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = e + 1;
    return true;
  }
};

static void tokens(const std::string& line, std::vector<std::string>& t) {
  t.clear();
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
    if (i >= line.size()) break;
    size_t st = i;
    while (i < line.size() && !std::isspace((unsigned char)line[i])) ++i;
    t.emplace_back(line, st, i - st);
  }
}

static bool to_id(const std::string& tok, int64_t base, uint32_t& out) {
  if (tok.empty()) return false;
  uint64_t v = 0;
  for (char c : tok) {
    if (c < '0' || c > '9') return false;
    v = v * 10 + (uint64_t)(c - '0');
    if (v > (uint64_t)INT32_MAX + 1) return false;
  }
  if ((int64_t)v < base) return false;
  out = (uint32_t)((int64_t)v - base);
  return true;
}

struct EdgeAcc {
  std::vector<uint32_t> s, d;
  uint64_t maxid = 0;
  bool any = false;
  void add(uint32_t a, uint32_t b) {
    s.push_back(a);
    d.push_back(b);
    maxid = std::max<uint64_t>(maxid, std::max(a, b));
    any = true;
  }
  int build(int symmetric, rg_hgraph** out) {
    return rg_hgraph_from_edges(any ? (int64_t)maxid + 1 : 0, (int64_t)s.size(), s.data(), d.data(), symmetric, out);
  }
};

// .wel: "u v w" triples (ReadInWEL, reader.h:59-67).
static int read_wel(const std::string& buf, int symmetric, rg_hgraph** out) {
  std::vector<std::pair<size_t, size_t>> w;
  split_words(buf, w);
  if (w.size() % 3) return set_error(RG_EFORMAT, "weighted edge list: not u v w triples");
  EdgeAcc el;
  for (size_t i = 0; i < w.size(); i += 3) {
    uint32_t a, b;
    if (!to_id(buf.substr(w[i].first, w[i].second), 0, a) || !to_id(buf.substr(w[i + 1].first, w[i + 1].second), 0, b))
      return set_error(RG_EFORMAT, "Bad input file");
    el.add(a, b);
  }
  return el.build(symmetric, out);
}

// .gr (DIMACS): "a u v w" arc lines, ids kept as written; other lines
// ignored (ReadInGR, reader.h:69-83).
static int read_gr(const std::string& buf, int symmetric, rg_hgraph** out) {
  TextLines lines(buf);
  std::string line;
  std::vector<std::string> t;
  EdgeAcc el;
  while (lines.next(line)) {
    if (line.empty() || line[0] != 'a') continue;
    tokens(line, t);
    uint32_t a, b;
    if (t.size() < 3 || t[0] != "a" || !to_id(t[1], 0, a) || !to_id(t[2], 0, b))
      return set_error(RG_EFORMAT, "Bad input file");
    el.add(a, b);
  }
  return el.build(symmetric, out);
}

// .graph (METIS): header "n m [fmt]" (fmt 1: weighted), then one line per
// vertex listing its 1-based neighbours (with weights for fmt 1); '%' lines
// are comments (ReadInMetis, reader.h:85-141).
static int read_metis(const std::string& buf, int symmetric, rg_hgraph** out) {
  TextLines lines(buf);
  std::string line;
  std::vector<std::string> t;
  bool header = false, weights = false;
  int64_t num_nodes = 0, u = 0;
  EdgeAcc el;
  while (lines.next(line)) {
    if (!line.empty() && line[0] == '%') continue;
    if (!header) {
      tokens(line, t);
      if (t.size() < 2) return set_error(RG_EFORMAT, "METIS header needs n and m");
      num_nodes = std::atoll(t[0].c_str());
      if (t.size() >= 3) {
        const int fmt = std::atoi(t[2].c_str());
        if (fmt == 1) weights = true;
        else if (fmt != 0 && fmt != 100)
          return set_error(RG_EFORMAT, "Do not support METIS fmt type: " + t[2]);
      }
      header = true;
      continue;
    }
    if (u >= num_nodes) break;
    tokens(line, t);
    const size_t step = weights ? 2 : 1;
    if (weights && t.size() % 2) return set_error(RG_EFORMAT, "METIS: neighbour without weight");
    for (size_t i = 0; i < t.size(); i += step) {
      uint32_t v;
      if (!to_id(t[i], 1, v)) return set_error(RG_EFORMAT, "Bad input file");
      el.add((uint32_t)u, v);
    }
    ++u;
  }
  if (!header) return set_error(RG_EFORMAT, "METIS header missing");
  return el.build(symmetric, out);
}

// .mtx: "%%MatrixMarket matrix coordinate <field> <symmetry>", '%'
// comments, "rows cols nonzeros", then 1-based "u v [w]" entries;
// symmetric matrices add (v, u) (ReadInMTX, reader.h:143-217).
static int read_mtx(const std::string& buf, int symmetric, rg_hgraph** out) {
  TextLines lines(buf);
  std::string line;
  std::vector<std::string> t;
  if (!lines.next(line)) return set_error(RG_EFORMAT, ".mtx file did not start with %%MatrixMarket");
  tokens(line, t);
  if (t.size() < 5 || t[0] != "%%MatrixMarket") return set_error(RG_EFORMAT, ".mtx file did not start with %%MatrixMarket");
  if (t[1] != "matrix" || t[2] != "coordinate")
    return set_error(RG_EFORMAT, "only allow matrix coordinate format for .mtx");
  if (t[3] == "complex") return set_error(RG_EFORMAT, "do not support complex weights for .mtx");
  bool weights;
  if (t[3] == "pattern") weights = false;
  else if (t[3] == "real" || t[3] == "double" || t[3] == "integer") weights = true;
  else return set_error(RG_EFORMAT, "unrecognized field type for .mtx");
  bool undirected;
  if (t[4] == "symmetric") undirected = true;
  else if (t[4] == "general" || t[4] == "skew-symmetric") undirected = false;
  else return set_error(RG_EFORMAT, "unsupported symmetry type for .mtx");
  bool sized = false;
  EdgeAcc el;
  while (lines.next(line)) {
    if (!line.empty() && line[0] == '%') continue;
    tokens(line, t);
    if (t.empty()) continue;
    if (!sized) {
      if (t.size() < 3 || t[0] != t[1]) return set_error(RG_EFORMAT, "matrix must be square for .mtx");
      sized = true;
      continue;
    }
    uint32_t a, b;
    if (t.size() < (weights ? 3u : 2u) || !to_id(t[0], 1, a) || !to_id(t[1], 1, b))
      return set_error(RG_EFORMAT, "Bad input file");
    el.add(a, b);
    if (undirected) el.add(b, a);
  }
  return el.build(symmetric, out);
}

static int read_text_format(const char* path, int format, int symmetric, rg_hgraph** out) {
  std::string buf;
  if (!read_file(path, buf)) return set_error(RG_EIO, std::string("cannot read ") + path);
  switch (format) {
    case RG_FMT_WEL: return read_wel(buf, symmetric, out);
    case RG_FMT_GR: return read_gr(buf, symmetric, out);
    case RG_FMT_METIS: return read_metis(buf, symmetric, out);
    default: return read_mtx(buf, symmetric, out);
  }
}

int rg_format_of(const char* path) {
  if (!path) return RG_FMT_ADJ;
  const std::string f(path);
  auto ends = [&](const char* suf) {
    const size_t k = std::strlen(suf);
    return f.size() > k && f.compare(f.size() - k, k, suf) == 0;
  };
  if (ends(".el")) return RG_FMT_EL;
  if (ends(".wel")) return RG_FMT_WEL;
  if (ends(".sg")) return RG_FMT_SG;
  if (ends(".gr")) return RG_FMT_GR;
  if (ends(".graph")) return RG_FMT_METIS;
  if (ends(".mtx")) return RG_FMT_MTX;
  return RG_FMT_ADJ;
}

int rg_hgraph_read(const char* path, int format, int symmetric, rg_hgraph** out) {
  if (!path || !out) return set_error(RG_EINVAL, "rg_hgraph_read: null argument");
  try {
    switch (format) {
      case RG_FMT_ADJ: return read_adj_text(path, symmetric, out);
      case RG_FMT_ADJ_BINARY: return read_adj_binary(path, symmetric, out);
      case RG_FMT_EL: return read_el(path, symmetric, out);
      case RG_FMT_SG: return read_sg(path, out);
      case RG_FMT_WEL:
      case RG_FMT_GR:
      case RG_FMT_METIS:
      case RG_FMT_MTX: return read_text_format(path, format, symmetric, out);
      default: return set_error(RG_EINVAL, "rg_hgraph_read: unknown format");
    }
  } catch (const std::bad_alloc&) {
    return set_error(RG_ENOMEM, "rg_hgraph_read: out of host memory");
  }
}

int rg_hgraph_write(const rg_hgraph* g, const char* path, int format) {
  if (!g || !path) return set_error(RG_EINVAL, "rg_hgraph_write: null argument");
  if (format == RG_FMT_ADJ_BINARY) {
    if (g->m > (int64_t)UINT32_MAX) return set_error(RG_EUNSUPPORTED, "binary .idx offsets are 32-bit");
    std::string p(path);
    FILE* f = std::fopen((p + ".config").c_str(), "wb");
    if (!f) return set_error(RG_EIO, "cannot write " + p + ".config");
    std::fprintf(f, "%lld", (long long)g->n);
    std::fclose(f);
    f = std::fopen((p + ".adj").c_str(), "wb");
    if (!f) return set_error(RG_EIO, "cannot write " + p + ".adj");
    std::fwrite(g->nbr.data(), 4, g->nbr.size(), f);
    std::fclose(f);
    f = std::fopen((p + ".idx").c_str(), "wb");
    if (!f) return set_error(RG_EIO, "cannot write " + p + ".idx");
    std::vector<uint32_t> o32(g->n);
    for (int64_t i = 0; i < g->n; ++i) o32[i] = (uint32_t)g->off[i];
    std::fwrite(o32.data(), 4, o32.size(), f);
    std::fclose(f);
    return RG_OK;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(RG_EIO, std::string("cannot write ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  if (format == RG_FMT_ADJ) {
    std::fprintf(f, "AdjacencyGraph\n%lld\n%lld\n", (long long)g->n, (long long)g->m);
    for (int64_t i = 0; i < g->n; ++i) std::fprintf(f, "%llu\n", (unsigned long long)g->off[i]);
    for (int64_t j = 0; j < g->m; ++j) std::fprintf(f, "%u\n", g->nbr[j]);
  } else if (format == RG_FMT_EL) {
    for (int64_t u = 0; u < g->n; ++u)
      for (uint64_t j = g->off[u]; j < g->off[u + 1]; ++j) std::fprintf(f, "%lld %u\n", (long long)u, g->nbr[j]);
  } else if (format == RG_FMT_SG) {
    char directed = g->symmetric ? 0 : 1;
    int64_t m = g->m, n = g->n;
    std::fwrite(&directed, 1, 1, f);
    std::fwrite(&m, 8, 1, f);
    std::fwrite(&n, 8, 1, f);
    std::fwrite(g->off.data(), 8, n + 1, f);
    std::fwrite(g->nbr.data(), 4, m, f);
    if (!g->symmetric) {
      std::fwrite(g->in_off.data(), 8, n + 1, f);
      std::fwrite(g->in_nbr.data(), 4, m, f);
    }
  } else {
    return set_error(RG_EINVAL, "rg_hgraph_write: unknown format");
  }
  return RG_OK;
}

// ---- source pickers -----------------------------------------------------

int rg_pick_sources_gap(const rg_hgraph* g, int count, uint32_t* out) {
  if (!g || !out || count < 0) return set_error(RG_EINVAL, "rg_pick_sources_gap: bad arguments");
  if (g->m == 0) return set_error(RG_EINVAL, "rg_pick_sources_gap: graph has no edges");
  std::mt19937 rng(rg::kRandSeed);
  std::uniform_int_distribution<int32_t> udist(0, (int32_t)(g->n - 1));
  for (int i = 0; i < count; ++i) {
    int32_t s;
    do {
      s = udist(rng);
    } while (g->off[s + 1] == g->off[s]);
    out[i] = (uint32_t)s;
  }
  return RG_OK;
}

// BC's source sequence (BC-manyiter-v2.C:149-168): mt19937(27491095) with
// uniform_int_distribution<long>(0, n-1), rejecting sources whose out-degree
// is 0 in the graph's *current* orientation.  BC transposes a directed graph
// once per source and never undoes it (BC-manyiter-v2.C:217), so source i is
// tested against out-degrees for even i and in-degrees for odd i.  Ids are in
// the graph's original space (a degree-sorted run maps them through new_ids,
// BC-manyiter-v2.C:176-177; the degree test is the same either way).
int rg_pick_sources_bc(const rg_hgraph* g, int count, uint32_t* out) {
  if (!g || !out || count < 0) return set_error(RG_EINVAL, "rg_pick_sources_bc: bad arguments");
  if (g->m == 0) return set_error(RG_EINVAL, "rg_pick_sources_bc: graph has no edges");
  std::mt19937 rng(rg::kRandSeed);
  std::uniform_int_distribution<long> udist(0, (long)g->n - 1);
  for (int i = 0; i < count; ++i) {
    const std::vector<uint64_t>& off = (!g->symmetric && (i & 1)) ? g->in_off : g->off;
    long s;
    do {
      s = udist(rng);
    } while (off[s + 1] == off[s]);
    out[i] = (uint32_t)s;
  }
  return RG_OK;
}

// Ligra's 64-bit integer hash (utils.h:347-355); the constants and shifts
// must be the reference's exactly so Radii picks the same 64 sources.
uint64_t rg_hash_int(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dULL) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddULL) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ULL) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ULL) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cULL) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bULL) ^ (a >> 32);
  return a;
}

}  // extern "C"
