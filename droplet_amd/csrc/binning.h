// Tile binning by two stable bucketings (rows, then columns); see binning.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dl_internal.h"

namespace dl {

// LDS-bounded grid limits of the span path (larger grids use emit + sort)
constexpr uint32_t kMaxBinTilesX = 256;
constexpr uint32_t kMaxBinTilesY = 1024;

struct BinGrid {
  uint32_t tiles_x, tiles_y, max_cov;
};

// Per-row prefix counts of the tiles that are still open: pre[y * stride +
// x] = open tiles of (virtual) row y in columns < x; stride = columns + 1.
// A null map means every tile is open.
struct LiveMap {
  const uint16_t *pre = nullptr;
  uint32_t stride = 0;
};

// Level-1 geometry of one view. A batch stacks the tile rows of its views
// into one virtual image (rows [row0, row0 + rows) belong to this view);
// items [item0, item0 + n) of the depth-ordered rect array are the view's,
// level-1 blocks [blk0, blk0 + ceil(n / 256)) process them, and their
// per-row counts start at c1_base.
struct L1View {
  uint32_t item0, n, blk0, row0, rows, c1_base;
  uint32_t tiles_x, max_cov;  // the view's tile grid is tiles_x x rows
};

// Packed rects of every view of a batch: rects[item0 + j] for the j-th
// depth-ordered item (ids[item0 + j], an index of rec) of each view; blocks
// of 256 items as the level-1 blocks.
void batch_rects(hipStream_t s, const float4 *rec, const uint32_t *ids, const L1View *tab, uint32_t ntab,
                 uint32_t blocks, uint4 *rects);

struct BinWorkspace {
  // level 1
  DeviceBuffer rects, c1, base1, spans, row_start;
  // level 2a / 2b chunk tables, counts and scanned bases
  DeviceBuffer tmp, chunk_base_a, seg_of_a, c2, base2, gspans;
  DeviceBuffer seg_start_b, chunk_base_b, seg_of_b, c3, base3;
  DeviceBuffer scan_ws;
};

bool span_binning_supported(uint32_t tiles_x, uint32_t tiles_y);

// Inputs of span_bin_rects: depth-ordered packed rects (uint4: item id, x0
// | x1 << 16, y0 | full_rows << 16, rem; rows local to the item's view)
// binned into a virtual image of rows x cols tiles.
struct SpanBinArgs {
  const uint4 *rects = nullptr;
  L1View one{};                  // the single view (tab == nullptr)
  const L1View *tab = nullptr;   // or a device table of ntab views, ordered by blk0 / row0
  uint32_t ntab = 0;
  uint32_t l1_blocks = 0;        // level-1 blocks over all views
  size_t n1 = 0;                 // level-1 counts over all views
  uint32_t rows = 0, cols = 0;   // virtual image, in tiles (cols <= kMaxBinTilesX)
  uint32_t max_view_rows = 0;    // LDS rows of a level-1 block (<= kMaxBinTilesY)
  uint32_t pair_cap = 0;         // capacity of the span, group-span and pvals arrays
  LiveMap live;                  // only open tiles receive entries
};

// Per-tile lists of every virtual tile: d_off[0..rows*cols] (lower-bound
// offsets, clamped to pair_cap) and pvals (item ids in depth order inside a
// tile). Runs entirely on the device. d_small holds three u32 device
// scalars: spans, group spans and list entries.
void span_bin_rects(hipStream_t s, const SpanBinArgs &a, BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals,
                    uint32_t *d_small);

// One view: ids[0..cn) are the chunk's splat ids in depth order, rec the
// projected records indexed by splat id; npairs bounds the chunk's (tile,
// splat) pairs.
void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              const LiveMap &live, BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals, uint32_t *d_small);

}  // namespace dl
