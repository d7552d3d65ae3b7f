// The voxel grid resident in HBM, shared by the voxelizer (voxel.hip) and
// the collision stages (collision.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "dl_internal.h"

namespace dl {

// One BlockMaskBuffer (src/voxel/block-mask-buffer.ts:23-58) on the device:
// the dense grid of 64-bit block masks, stored as uint2 (lo, hi) at the flat
// block index bx + by*nb0 + bz*nb0*nb1 (voxel bit x + 4y + 16z), and the
// non-empty blocks in the voxelizer's insertion order. Host copies of the
// lists are made on first request.
struct VoxelResult {
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  double nblocks[3] = {0, 0, 0};  // VoxelizeResult.numBlocksX/Y/Z (JS numbers)
  double vres = 0.05;             // voxel side length
  int32_t nb[3] = {0, 0, 0};      // blocks per axis; 0 for an empty grid
  uint64_t total = 0;             // nb0 * nb1 * nb2
  uint64_t count = 0;             // non-empty blocks
  int device = 0;
  DeviceBuffer masks;             // uint2[total]
  DeviceBuffer idx, lo, hi;       // u32[count] each, insertion order
  mutable std::vector<uint32_t> h_idx, h_lo, h_hi;
  mutable bool host_ready = false;
  float device_ms = 0.0f;
  float eval_ms = 0.0f;           // the per-voxel Gaussian sums (k_vox_eval)
  uint64_t block_pairs = 0;       // (block, Gaussian) pairs the sums read
};

// Builds r.idx / r.lo / r.hi and r.count from r.masks (r.nb and r.total
// set), with the context's compaction scratch. Records `done` (if not null)
// after the last kernel, then waits for the stream.
void voxel_compact(Context &c, hipStream_t s, VoxelResult &r, hipEvent_t done);

// Copies the non-empty block lists to the host vectors once.
void voxel_lists_to_host(const VoxelResult &r);

}  // namespace dl

struct dl_voxel_result {
  dl::VoxelResult r;
};
