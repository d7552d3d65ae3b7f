// Tile binning by two stable bucketings (rows, then columns); see binning.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dl_common.h"
#include "dl_internal.h"

namespace dl {

// Splats per level-1 block (one per thread). 128 and 512 bin a 6M-splat 4K
// frame 13 us and 3 us slower. The batch's level-1 tables count blocks of
// this size.
constexpr int kB1 = 256;

// LDS-bounded grid limits of the span path (larger grids use emit + sort)
constexpr uint32_t kMaxBinTilesX = 256;
constexpr uint32_t kMaxBinTilesY = 1024;

struct Rect {
  uint32_t x0, x1, y0;
  uint32_t nrows;               // rows y0 .. y0+nrows-1 (0: no pairs)
  uint32_t full_rows;           // rows y0 .. y0+full_rows-1 span [x0, x1]
  uint32_t rem;                 // row y0+full_rows spans [x0, x0+rem-1] when rem > 0
};

// tile rectangle and maxCoverage truncation of one projected splat, with the
// exact arithmetic of project.ts:223-233 / emit-pairs.ts:38-41 at 16-px
// tiles (tile_range gives the range at the autotuner's 8- and 32-px tiles)
__device__ inline Rect splat_rect(const float4 v0, const TileGrid &g, uint32_t max_cov) {
  Rect R{0, 0, 0, 0, 0, 0};
  const float sx = v0.x, sy = v0.y, radius = v0.z;
  if (!(radius > 0.0f)) return R;
  uint32_t min_tx, max_tx, min_ty, max_ty;
  if (!tile_range(sx, sy, radius, g, min_tx, max_tx, min_ty, max_ty)) return R;
  const uint32_t w = max_tx - min_tx + 1u, h = max_ty - min_ty + 1u;
  R.x0 = min_tx;
  R.x1 = max_tx;
  R.y0 = min_ty;
  if ((uint64_t)w * h <= max_cov) {
    R.full_rows = h;
    R.rem = 0;
  } else {
    R.full_rows = max_cov / w;
    R.rem = max_cov - R.full_rows * w;
  }
  R.nrows = R.full_rows + (R.rem > 0 ? 1u : 0u);
  return R;
}

// the grid of a 16-px tiles_x x tiles_y image
__host__ __device__ inline TileGrid grid16(uint32_t tiles_x, uint32_t tiles_y) {
  return TileGrid{kTile, tiles_x, tiles_y, tiles_x, tiles_y};
}

// Rect of a span-binnable grid (x < 256, y < 1024, rem < 256) in 8 bytes:
// (x0 | x1 << 8 | rem << 16, y0 | full_rows << 16)
__device__ inline uint2 pack_rect8(const Rect &R) {
  return make_uint2(R.x0 | (R.x1 << 8) | (R.rem << 16), R.y0 | (R.full_rows << 16));
}

__device__ inline Rect unpack_rect8(const uint2 p) {
  Rect R;
  R.x0 = p.x & 0xFFu;
  R.x1 = (p.x >> 8) & 0xFFu;
  R.rem = p.x >> 16;
  R.y0 = p.y & 0xFFFFu;
  R.full_rows = p.y >> 16;
  R.nrows = R.full_rows + (R.rem > 0 ? 1u : 0u);
  return R;
}

// binning grid: tiles_x x tiles_y tiles of tg.tile pixels (tg's grid);
// max_cov is maxCoverage in those tiles
struct BinGrid {
  uint32_t tiles_x, tiles_y, max_cov;
  TileGrid tg;
};

// Level 1 reads each splat as one 16-byte record: id, then the rect.
__device__ inline uint4 pack_rect(uint32_t id, const Rect &R) {
  return make_uint4(id, R.x0 | (R.x1 << 16), R.y0 | (R.full_rows << 16), R.rem);
}

// Per-row prefix counts of the tiles that are still open: pre[y * stride +
// x] = open tiles of (virtual) row y in columns < x; stride = columns + 1.
// A null map means every tile is open.
struct LiveMap {
  const uint16_t *pre = nullptr;
  uint32_t stride = 0;
  // summed-area table of the open tiles, same stride, rows + 1 rows (set by
  // the live-splat compaction)
  const uint32_t *sat = nullptr;
  // the same table at the corners of 4x4-tile blocks: csat[cy * cstride +
  // cx] = sat[min(4 cy, rows) * stride + min(4 cx, cols)], small enough to
  // stay in a CU's L1
  const uint32_t *csat = nullptr;
  uint32_t cstride = 0;
};

// Does row y (virtual) hold a live tile in columns [xa, xb]? Without a map
// every tile is live.
__device__ inline bool live_range(const LiveMap &L, uint32_t y, uint32_t xa, uint32_t xb) {
  if (!L.pre) return true;
  const uint16_t *r = L.pre + (size_t)y * L.stride;
  return r[xb + 1] != r[xa];
}

// Does a splat's rect (rows from y0 of the live map's virtual image) hold an
// open tile: in its full rows (the summed-area table) or in its partial last
// row (the row prefix)? The 4x4-block table first: no open tile in the
// blocks that cover the rect (rem row included) means none in the rect.
__device__ inline bool rect_live(const Rect &R, const LiveMap &live) {
  if (R.nrows == 0) return false;
  const uint32_t st = live.stride, cs = live.cstride;
  const uint32_t cy0 = R.y0 >> 2, cy1 = (R.y0 + R.nrows + 3u) >> 2, cx0 = R.x0 >> 2, cx1 = (R.x1 + 4u) >> 2;
  const bool maybe = live.csat[cy1 * cs + cx1] - live.csat[cy0 * cs + cx1] - live.csat[cy1 * cs + cx0] +
                         live.csat[cy0 * cs + cx0] != 0u;
  bool lv = false;
  if (maybe && R.full_rows > 0) {
    const uint32_t ya = R.y0, yb = R.y0 + R.full_rows, xa = R.x0, xb = R.x1 + 1;
    const uint32_t open = live.sat[(size_t)yb * st + xb] - live.sat[(size_t)ya * st + xb] -
                          live.sat[(size_t)yb * st + xa] + live.sat[(size_t)ya * st + xa];
    lv = open != 0;
  }
  if (maybe && !lv && R.rem > 0) lv = live_range(live, R.y0 + R.full_rows, R.x0, R.x0 + R.rem - 1u);
  return lv;
}

// Level-1 geometry of one view. A batch stacks the tile rows of its views
// into one virtual image (rows [row0, row0 + rows) belong to this view);
// items [item0, item0 + n) of the depth-ordered rect array are the view's,
// level-1 blocks [blk0, blk0 + ceil(n / 256)) process them, and their
// per-row counts start at c1_base.
struct L1View {
  uint32_t item0, n, blk0, row0, rows, c1_base;
  uint32_t tiles_x, max_cov;  // the view's tile grid is tiles_x x rows
  // items level 1 reads (<= n; the layout of blocks and counts follows n):
  // n, or the live splats that a batch pass compacted to the front
  uint32_t n_live;
  // a batch view's first item in the group's item arrays and first word of
  // its live bits (compacted passes)
  uint32_t vitem0, bit0;
};

// Packed rects of every view of a batch: rects[item0 + j] for the j-th
// depth-ordered item (ids[item0 + j], an index of rec) of each view; blocks
// of 256 items as the level-1 blocks.
void batch_rects(hipStream_t s, const float4 *rec, const uint32_t *ids, const L1View *tab, uint32_t ntab,
                 uint32_t blocks, uint4 *rects);

// The open tiles' summed-area tables of a tall stacked image (rows x cols
// tiles, row prefixes pre): sat ((rows + 1) x (cols + 1)) and csat
// ((ceil(rows / 4) + 1) x (ceil(cols / 4) + 1)), one workgroup per column.
LiveMap open_sat_tall(hipStream_t s, const uint16_t *pre, uint32_t rows, uint32_t cols, uint32_t *sat,
                      uint32_t *csat);

struct BinWorkspace {
  // level 1
  DeviceBuffer rects, c1, base1, spans, row_start;
  // level 2a / 2b chunk tables, counts and scanned bases
  DeviceBuffer tmp, chunk_base_a, seg_of_a, c2, base2, gspans;
  DeviceBuffer seg_start_b, chunk_base_b, seg_of_b, c3, base3;
  DeviceBuffer scan_ws;
  // live-splat compaction of later passes: id-order live bits, per-block
  // counts and offsets, [0] live items and [1] their level-1 blocks
  DeviceBuffer live_bits, live_cnt, live_small, live_sat, live_csat;
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
  // single view only: the item count on the device (items of `one` are then
  // min(*d_items, one.n)) and ceil(items / 256); null: one.n items
  const uint32_t *d_items = nullptr;
  const uint32_t *d_nb1 = nullptr;
  // when set, the pass's list entries are added to *entries on the device
  unsigned long long *entries = nullptr;
};

// Per-tile lists of every virtual tile: d_off[0..rows*cols] (lower-bound
// offsets, clamped to pair_cap) and pvals (item ids in depth order inside a
// tile). Runs entirely on the device. d_small holds three u32 device
// scalars: spans, group spans and list entries.
void span_bin_rects(hipStream_t s, const SpanBinArgs &a, BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals,
                    uint32_t *d_small);

// One view: ids[0..cn) are the chunk's splat ids in depth order, rec the
// projected records indexed by splat id; npairs bounds the chunk's (tile,
// splat) pairs. With a live map and rect8 (the packed rects of all n_all
// splats in id order), only splats with a live tile in some row of their
// rect enter level 1 (a stable compaction); the lists are the same.
void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              const LiveMap &live, const uint2 *rect8, uint32_t n_all, BinWorkspace &ws, uint32_t *d_off,
              uint32_t *pvals, uint32_t *d_small, unsigned long long *d_entries = nullptr,
              bool rects_ready = false);

}  // namespace dl
