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

struct BinWorkspace {
  // level 1
  DeviceBuffer c1, base1, spans, row_start;
  // level 2a / 2b chunk tables, counts and scanned bases
  DeviceBuffer tmp, chunk_base_a, seg_of_a, c2, base2, gspans;
  DeviceBuffer seg_start_b, chunk_base_b, seg_of_b, c3, base3;
  DeviceBuffer scan_ws;
};

bool span_binning_supported(uint32_t tiles_x, uint32_t tiles_y);

// Per-tile splat lists of one chunk: d_off[0..num_tiles] (lower-bound
// offsets) and pvals[0..npairs) (splat ids; inside a tile in chunk order,
// i.e. depth order). ids[0..cn) are the chunk's splat ids in depth order and
// rec holds the projected records indexed by splat id. npairs is the scanned total of the project() tile counts. d_small
// holds three u32 device scalars.
void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals, uint32_t *d_small);

}  // namespace dl
