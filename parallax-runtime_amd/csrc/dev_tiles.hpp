// dev_tiles.hpp -- tile geometry shared by the tiled (scan, compaction, segmented
// reduce) kernels.
//
// Every tiled kernel reads through 16-byte vector loads from the 16-byte-aligned
// address at or below the caller's pointer (`shift` elements before it).  A segment
// [off[s], off[s+1]) of the caller's array is the range [shift+off[s], shift+off[s+1])
// in those aligned coordinates; its tiles are consecutive windows of TILE elements
// starting at the aligned-down segment start, so every vector load is aligned no matter
// where a segment begins.  Positions outside the segment are masked to the identity and
// never stored.  The plain (one-segment) form uses the same geometry with no tables,
// which is what makes a segment's result identical alone and inside a batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.hpp"

namespace pxk {

struct SegGeom {
  // Segment s is [seg_lo[s], seg_hi[s]) of the input (element indices relative to the
  // caller's input pointer); its outputs start at seg_out[s] of the output pointer
  // (seg_out == null: at seg_lo[s]).  seg_lo == null: one segment [0, n).
  const uint64_t* seg_lo;
  const uint64_t* seg_hi;
  const uint64_t* seg_out;
  const uint32_t* tile_off;  // nseg+1 first tile of each segment (null for one segment)
  const uint32_t* tile_seg;  // ntiles: segment of each tile (null for one segment)
  uint64_t n;                // element count when seg_lo is null
  uint32_t ntiles;           // total tiles of the launch
  uint32_t nseg;             // 1 when seg_lo is null
  int tile_block;            // tile shape: threads x 16-byte vectors per thread
  int tile_k;
};

struct TileInfo {
  uint32_t seg;
  uint32_t first_tile;  // first tile of this tile's segment
  uint32_t last_tile;   // last tile of this tile's segment
  int64_t lo, hi;       // valid positions [lo, hi) in aligned coordinates
  int64_t w0;           // window start (aligned coordinates, multiple of VEC)
  int64_t out_delta;    // output element index = input element index + out_delta
};

template <int VEC, int TILE>
__device__ __forceinline__ TileInfo tile_info(const SegGeom& g, uint32_t tile, int64_t shift) {
  TileInfo t;
  tile = __builtin_amdgcn_readfirstlane(tile);
  if (g.seg_lo == nullptr) {
    t.seg = 0;
    t.first_tile = 0;
    t.last_tile = g.ntiles - 1;
    t.lo = shift;
    t.hi = shift + static_cast<int64_t>(g.n);
    t.w0 = static_cast<int64_t>(tile) * TILE;
    t.out_delta = 0;
  } else {
    t.seg = ld_table(g.tile_seg, tile);
    t.first_tile = ld_table(g.tile_off, t.seg);
    t.last_tile = ld_table(g.tile_off, t.seg + 1) - 1;
    const int64_t b = static_cast<int64_t>(ld_table(g.seg_lo, t.seg));
    t.lo = shift + b;
    t.hi = shift + static_cast<int64_t>(ld_table(g.seg_hi, t.seg));
    t.w0 = (t.lo & ~static_cast<int64_t>(VEC - 1)) +
           static_cast<int64_t>(tile - t.first_tile) * TILE;
    t.out_delta = g.seg_out ? static_cast<int64_t>(ld_table(g.seg_out, t.seg)) - b : 0;
  }
  return t;
}

// Element shift of a pointer below its 16-byte-aligned base.
__host__ __device__ inline int64_t elem_shift(const void* p, size_t elem_size) {
  return static_cast<int64_t>((reinterpret_cast<uintptr_t>(p) & 15u) / elem_size);
}

}  // namespace pxk
