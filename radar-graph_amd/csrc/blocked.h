// Destination-blocked in-edge layout shared by the blocked PageRank engine
// and the dense levels of the bit-parallel multi-source BFS.
//
// Destinations are cut into blocks of 2^vb vertices; the in-edges of a block
// are sorted by source (segment) and stored as a 4-byte source id plus a
// 2-byte local destination.  A block's edge range is split into units of at
// most kUnitEdges edges, one workgroup each; a block with several units (the
// hub blocks of a degree-sorted graph) gets one private LDS copy per unit,
// merged afterwards.
#pragma once

#include <cstdint>

#include "device.h"

namespace rg {

constexpr int64_t kUnitEdges = int64_t(1) << 19;
// Destination-block bits of the layout the query kernels share (batched
// BFS, dense BC levels): 4096 vertices, whose 64-bit words or FP64 values
// fit in LDS.
constexpr int kQueryVB = 12;

// Units per merge chunk (BlockedCSC::mchunk).
constexpr uint32_t kMergeChunk = 16;

struct BlockedCSC {
  const uint64_t* built_for = nullptr;   // in_off of the orientation it was built from
  int vb = 0, sb = 0;                    // destination-block / source-segment bits
  int64_t nblocks = 0, nunits = 0;
  int64_t nmulti = 0;                    // blocks with more than one unit
  uint32_t* src = nullptr;         // m, sorted by (block, segment, local dst)
  uint16_t* vl = nullptr;          // m, local destination
  uint64_t* unit_lo = nullptr;     // nunits + 1 edge boundaries
  uint32_t* block_unit0 = nullptr; // nblocks + 1: first unit of each block
  uint32_t* unit_block = nullptr;  // nunits: block of each unit
  uint32_t* multi = nullptr;       // nmulti: the blocks with more than one unit
  // Merge chunks: the units of every split block in runs of kMergeChunk
  // (the last run of a block shorter); mchunk[c] = the run's first unit.
  uint32_t* mchunk = nullptr;      // nmchunk
  int64_t nmchunk = 0;
  // Launch order other than the layout's (workgroup or queue position ->
  // unit): the split blocks' units sorted by their first source -- a split
  // hub block's units cover consecutive source ranges, so units of different
  // hub blocks that read the same sources then run at the same time --
  // followed by the single-unit blocks in block order.  nsplit = the split
  // blocks' unit count; xcd_shares (device, 2 (X + 1) entries) cuts the
  // sorted split units into X contiguous per-XCD shares [lo, hi) plus the
  // single-unit blocks as a common share, for k_blk_pr_qx.
  uint32_t* by_src = nullptr;      // nunits
  int64_t nsplit = 0;
  int64_t* xcd_shares = nullptr;
  int nxcd = 0;
  // Packed arcs (the PageRank layout, when 2^sbits-source segments leave
  // room): pk[e] = (source mod 2^sbits) << vb | local destination, 4 bytes
  // per arc in place of src + vl (which are then freed).  A unit's arcs,
  // sorted by source, fall in at most kMaxSegs segments of 2^sbits
  // sources: segment j of unit u (j in [unit_seg0[u], unit_seg0[u + 1]))
  // starts at arc unit_lo[u] + seg_rel[j] and holds sources seg_hi[j] <<
  // sbits + (pk >> vb).
  uint32_t* pk = nullptr;          // m
  uint32_t* seg_rel = nullptr;     // nsegs
  uint32_t* seg_hi = nullptr;      // nsegs
  uint32_t* unit_seg0 = nullptr;   // nunits + 1
  int64_t nsegs = 0;
  int sbits = 0;
  bool pack_asked = false;
  void release();
};

// Builds (or returns the cached) blocked layout of g's current in-edges in
// the cache slot *slot (a member of rg_graph).  The layout is read-only once
// built, so contexts sharing the graph may use it at the same time; the
// per-unit partial results live in each context's scratch.
// pack: build the packed arcs (BlockedCSC::pk) where the graph allows.
int blocked_build(rg_graph* g, int vb, int sb, void** slot, BlockedCSC** out, bool pack = false);

constexpr int kMaxSegs = 1024;   // source segments per unit in the packed layout

// The query layout (2^kQueryVB-vertex blocks, source order) of g's current
// orientation, built on first use.
inline int query_layout(rg_graph* g, BlockedCSC** out) {
  return blocked_build(g, kQueryVB, 0, g->transposed ? &g->blocked_ms_t : &g->blocked_ms, out);
}

}  // namespace rg
