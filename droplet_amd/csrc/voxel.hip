// Cooperative voxelizer on gfx950: the device path behind dl_voxelize.
//
// Reference (citations relative to the droplet repository):
//   computeAabbExtents        src/spatial/aabb.ts:22-62            (f64)
//   computeBlockAlignedBounds src/voxel/voxelize.ts:278-306        (f64)
//   batch loop / BVH query    src/voxel/voxelize.ts:185-234, src/spatial/bvh.ts:136-177
//   voxelizeShader            src/voxel/shaders/voxelize.ts:17-148
//   mask decode + addBlock    src/voxel/voxelize.ts:160-179, src/voxel/block-mask-buffer.ts:30-41
//
// Device design. The reference gives every 4x4x4 block of a 16^3-block batch
// the batch's whole BVH candidate list (~10^4 Gaussians) and tests all of
// them per voxel. Here every Gaussian is binned directly to the blocks it can
// reach: the block box is the intersection of (a) the batches whose f64 box
// its f64 AABB overlaps (the reference's exact batch-list test) and (b) a
// conservative f64 superset of the blocks holding a voxel centre that can
// pass the f32 per-voxel AABB test. Pairs (block, gaussian) are sorted stably
// by block, so each block's list is in Gaussian index order. One wave64 per
// block then evaluates its list exactly as evalGaussian does, 64 candidates
// at a time through LDS, with the sigma >= 7 early-out. Gaussians a block's
// list omits fail the per-voxel test for every voxel of that block and would
// add exactly +0.0, so every voxel sees the same non-zero terms as the
// reference; only their summation order differs (index order instead of BVH
// DFS order), which moves alpha by float rounding only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "dl_common.h"
#include "dl_internal.h"
#include "voxel_grid.h"

namespace dl {

namespace {

// Gaussian record of the voxel kernel (the reference uploads 16 floats per
// Gaussian, voxelize.ts:250-276); exp(-logScale) and the sigmoid opacity are
// precomputed with the same f32 functions evalGaussian applies.
struct VGauss {
  float px, py, pz, ex;
  float ey, ez, qw, qx;
  float qy, qz, is0, is1;
  float is2, opacity, pad0, pad1;
};
static_assert(sizeof(VGauss) == 64, "VGauss is 64 bytes");

struct VoxGrid {
  double bmn[3];        // grid bounds min (f64)
  double bs;            // block size (f64)
  int32_t nb[3];        // blocks per axis
  uint32_t nbatch[3];   // batches per axis
  float res, half, cutoff;
  float stop;           // sigma at which a voxel's sum ends (voxel_solid_stop)
  uint32_t layered;     // block lists keyed ((v nbu + u) << lb) | l (else by flat index)
  uint32_t lax;         // the layer axis l (0-2); u and v are the other two, u the lower
  uint32_t lb;          // bits of the layer index l in those keys
};

// A Gaussian's block box on the layer route's axes: layers [l0, l1] of the
// layer axis, columns [u0, u1] and rows [v0, v1] of the other two.
struct LayerBox {
  int32_t l0, l1, u0, u1, v0, v1;
};

__device__ inline LayerBox layer_box(const VoxGrid &G, const int32_t lo[3], const int32_t hi[3]) {
  const uint32_t L = G.lax;
  LayerBox b;
  b.l0 = L == 0 ? lo[0] : (L == 1 ? lo[1] : lo[2]);
  b.l1 = L == 0 ? hi[0] : (L == 1 ? hi[1] : hi[2]);
  b.u0 = L == 0 ? lo[1] : lo[0];
  b.u1 = L == 0 ? hi[1] : hi[0];
  b.v0 = L == 2 ? lo[1] : lo[2];
  b.v1 = L == 2 ? hi[1] : hi[2];
  return b;
}

// blocks along the u axis, and along the layer axis
__device__ inline uint32_t layer_nbu(const VoxGrid &G) { return (uint32_t)(G.lax == 0 ? G.nb[1] : G.nb[0]); }
__device__ inline uint32_t layer_nbl(const VoxGrid &G) {
  return (uint32_t)(G.lax == 0 ? G.nb[0] : (G.lax == 1 ? G.nb[1] : G.nb[2]));
}

// the layer route's key of block (ax, ay, az)
__device__ inline uint32_t layer_key(const VoxGrid &G, uint32_t ax, uint32_t ay, uint32_t az) {
  const uint32_t L = G.lax;
  const uint32_t l = L == 0 ? ax : (L == 1 ? ay : az);
  const uint32_t u = L == 0 ? ay : ax;
  const uint32_t v = L == 2 ? ay : az;
  return ((v * layer_nbu(G) + u) << G.lb) | l;
}

constexpr int kPrepBlock = 256;

// computeAabbExtents (aabb.ts:22-62) in f64 + the record + bounds partials
__global__ __launch_bounds__(kPrepBlock) void k_vox_prep(const float *__restrict__ pos, const float *__restrict__ rot,
                                                         const float *__restrict__ ls, const float *__restrict__ op,
                                                         uint32_t n, VGauss *__restrict__ rec,
                                                         double *__restrict__ partials) {
  __shared__ double s_red[6][kPrepBlock / 64];
  const uint32_t i = blockIdx.x * kPrepBlock + threadIdx.x;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    const double qw = rot[4 * (size_t)i], qx = rot[4 * (size_t)i + 1];
    const double qy = rot[4 * (size_t)i + 2], qz = rot[4 * (size_t)i + 3];
    const double sx = exp((double)ls[3 * (size_t)i]);
    const double sy = exp((double)ls[3 * (size_t)i + 1]);
    const double sz = exp((double)ls[3 * (size_t)i + 2]);
    const double xx = qx * qx, yy = qy * qy, zz = qz * qz;
    const double xy = qx * qy, xz = qx * qz, yz = qy * qz;
    const double wx = qw * qx, wy = qw * qy, wz = qw * qz;
    const double r00 = 1 - 2 * (yy + zz), r01 = 2 * (xy - wz), r02 = 2 * (xz + wy);
    const double r10 = 2 * (xy + wz), r11 = 1 - 2 * (xx + zz), r12 = 2 * (yz - wx);
    const double r20 = 2 * (xz - wy), r21 = 2 * (yz + wx), r22 = 1 - 2 * (xx + yy);
    double a, b, c;
    a = r00 * sx; b = r01 * sy; c = r02 * sz;
    const float ex = (float)(__builtin_sqrt((a * a + b * b) + c * c) * 3.0);
    a = r10 * sx; b = r11 * sy; c = r12 * sz;
    const float ey = (float)(__builtin_sqrt((a * a + b * b) + c * c) * 3.0);
    a = r20 * sx; b = r21 * sy; c = r22 * sz;
    const float ez = (float)(__builtin_sqrt((a * a + b * b) + c * c) * 3.0);
    VGauss g;
    g.px = pos[3 * (size_t)i];
    g.py = pos[3 * (size_t)i + 1];
    g.pz = pos[3 * (size_t)i + 2];
    g.ex = ex;
    g.ey = ey;
    g.ez = ez;
    g.qw = rot[4 * (size_t)i];
    g.qx = rot[4 * (size_t)i + 1];
    g.qy = rot[4 * (size_t)i + 2];
    g.qz = rot[4 * (size_t)i + 3];
    g.is0 = dl_expf(-ls[3 * (size_t)i]);
    g.is1 = dl_expf(-ls[3 * (size_t)i + 1]);
    g.is2 = dl_expf(-ls[3 * (size_t)i + 2]);
    g.opacity = 1.0f / (1.0f + dl_expf(-op[i]));
    g.pad0 = 0.0f;
    g.pad1 = 0.0f;
    rec[i] = g;
    const float p3[3] = {g.px, g.py, g.pz}, e3[3] = {ex, ey, ez};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = (double)p3[k] - (double)e3[k];
      mx[k] = (double)p3[k] + (double)e3[k];
    }
  }
  // block min/max (f64 min/max are exact, order-independent)
#pragma unroll
  for (int k = 0; k < 3; k++) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fmin(mn[k], __shfl_xor(mn[k], d, 64));
      mx[k] = fmax(mx[k], __shfl_xor(mx[k], d, 64));
    }
  }
  const int wave = threadIdx.x >> 6;
  if (__lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_red[k][wave] = mn[k];
      s_red[3 + k][wave] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double v = s_red[k][0];
    for (int w = 1; w < kPrepBlock / 64; w++) v = k < 3 ? fmin(v, s_red[k][w]) : fmax(v, s_red[k][w]);
    partials[6 * (size_t)blockIdx.x + k] = v;
  }
}

// The bounds of all Gaussians from the per-workgroup partials of
// k_vox_prep: out[0..2] min, out[3..5] max (f64 min/max are exact; NaNs are
// passed over as by the host's std::min / std::max from +-infinity).
__global__ __launch_bounds__(1024) void k_vox_bounds(const double *__restrict__ partials, uint32_t nparts,
                                                     double *__restrict__ out) {
  __shared__ double s_red[6][16];
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[k] = INFINITY;
    v[3 + k] = -INFINITY;
  }
  for (uint32_t b = threadIdx.x; b < nparts; b += 1024) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = fmin(v[k], partials[6 * (size_t)b + k]);
      v[3 + k] = fmax(v[3 + k], partials[6 * (size_t)b + 3 + k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++)
    for (int d = 32; d >= 1; d >>= 1) {
      const double o = __shfl_xor(v[k], d, 64);
      v[k] = k < 3 ? fmin(v[k], o) : fmax(v[k], o);
    }
  if (__lane_id() == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) s_red[k][threadIdx.x >> 6] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double r = s_red[k][0];
    for (int w = 1; w < 16; w++) r = k < 3 ? fmin(r, s_red[k][w]) : fmax(r, s_red[k][w]);
    out[k] = r;
  }
}

// Per axis: blocks [lo, hi] (inclusive; lo > hi = none) that Gaussian with
// centre p and f32 extent e can reach under both filters.
__device__ inline void axis_block_range(const VoxGrid &G, int k, float pf, float ef, int32_t *lo, int32_t *hi) {
  const double p = pf, e = ef;
  const double gmn = p - e, gmx = p + e;
  const double bs = G.bs, bmn = G.bmn[k];
  const int32_t nb = G.nb[k];
  // (a) batches whose f64 box the f64 AABB overlaps: the exact test of
  // bvh.ts:158-171 against the batch box built as voxelize.ts:193-198
  const double span = 16.0 * bs;
  int32_t kb_lo = (int32_t)floor((gmn - bmn) / span) - 1;
  int32_t kb_hi = (int32_t)floor((gmx - bmn) / span) + 1;
  kb_lo = max(kb_lo, 0);
  kb_hi = min(kb_hi, (int32_t)G.nbatch[k] - 1);
  int32_t first = -1, last = -2;
  for (int32_t kb = kb_lo; kb <= kb_hi; kb++) {
    const int32_t b0 = kb * 16;
    const int32_t cnt = min(16, nb - b0);
    const double qmn = bmn + (double)b0 * bs;
    const double qmx = qmn + (double)cnt * bs;
    const bool overlap = !(gmx < qmn || gmn > qmx);
    if (overlap) {
      if (first < 0) first = kb;
      last = kb;
    }
  }
  if (first < 0) {
    *lo = 1;
    *hi = 0;
    return;
  }
  // (b) blocks whose voxel centres (block min + 0.5 res .. + 3.5 res) come
  // within e + half of p, widened by a margin far above f32 rounding
  const double reach = e + (double)G.half;
  const double res = (double)G.res;
  const double margin = 1e-5 * (fabs(bmn) + fabs(p) + reach + (double)nb * bs + 1.0);
  const double lo_f = (p - reach - margin - bmn - 3.5 * res) / bs;
  const double hi_f = (p + reach + margin - bmn - 0.5 * res) / bs;
  int32_t b_lo = (int32_t)fmax(ceil(lo_f), -1.0);
  int32_t b_hi = (int32_t)fmin(floor(hi_f), (double)nb);
  b_lo = max(b_lo, first * 16);
  b_hi = min(b_hi, min(last * 16 + 15, nb - 1));
  b_lo = max(b_lo, 0);
  *lo = b_lo;
  *hi = b_hi;
}

constexpr uint32_t kTotalSlots = 64;

// Inclusive scan of v over the wave's lanes with DPP row shifts and row
// broadcasts (register moves, no LDS round trip per step as the shuffles
// take): within each 16-lane row by shifts of 1, 2, 4 and 8, then row 1 and
// 3 add lane 15 of the row before, rows 2 and 3 lane 31.
__device__ inline uint32_t wave_inclusive_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
  return v;
}

// Workgroups are dealt to the 8 XCDs in turn. A launch of 8 q workgroups
// gives workgroup b the chunk (b mod 8) q + b / 8, so each XCD takes a
// contiguous run of chunks: the layer-route kernels' writes of neighbouring
// waves (adjacent in every layer) then meet in one L2.
__device__ inline uint32_t xcd_chunk() {
  const uint32_t q = gridDim.x >> 3;
  return (blockIdx.x & 7u) * q + (blockIdx.x >> 3);
}

// Least z0 and greatest z1 of a wave's Gaussians with blocks (zmin > zmax
// when none has any).
// (DPP steps as wave_inclusive_scan_dpp, with min and max; lane 63 then
// holds the wave's result, read into a scalar register.)
__device__ inline void wave_layer_range(bool any, int32_t z0, int32_t z1, int32_t *zmin, int32_t *zmax) {
  int32_t a = any ? z0 : 0x7FFFFFFF, b = any ? z1 : -1;
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x111, 0xf, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x111, 0xf, 0xf, false));
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x112, 0xf, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x112, 0xf, 0xf, false));
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x114, 0xf, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x114, 0xf, 0xf, false));
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x118, 0xf, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x118, 0xf, 0xf, false));
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x142, 0xa, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x142, 0xa, 0xf, false));
  a = min(a, __builtin_amdgcn_update_dpp(0x7FFFFFFF, a, 0x143, 0xc, 0xf, false));
  b = max(b, __builtin_amdgcn_update_dpp(-1, b, 0x143, 0xc, 0xf, false));
  *zmin = __builtin_amdgcn_readlane(a, 63);
  *zmax = __builtin_amdgcn_readlane(b, 63);
}

// Per Gaussian: its block count (counts, flat route). With zc (layer route)
// instead, wave w (Gaussians 64 w .. 64 w + 63) writes zc[z * nw + w], the
// pairs its Gaussians have in block layer z, for every layer z of the layer
// axis: the counts of the layer-major placement that k_vox_emit_layers
// writes.
__global__ __launch_bounds__(256) void k_vox_count(const VGauss *__restrict__ rec, uint32_t n, VoxGrid G,
                                                   uint32_t *__restrict__ counts,
                                                   unsigned long long *__restrict__ d_total64,
                                                   uint32_t *__restrict__ zc, uint32_t nw) {
  const uint32_t i = (zc ? xcd_chunk() : blockIdx.x) * 256 + threadIdx.x;
  unsigned long long c = 0;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  if (i < n) {
    const VGauss g = rec[i];
    axis_block_range(G, 0, g.px, g.ex, &lo[0], &hi[0]);
    axis_block_range(G, 1, g.py, g.ey, &lo[1], &hi[1]);
    axis_block_range(G, 2, g.pz, g.ez, &lo[2], &hi[2]);
    c = 1;
#pragma unroll
    for (int k = 0; k < 3; k++) c *= (hi[k] >= lo[k]) ? (unsigned long long)(hi[k] - lo[k] + 1) : 0ull;
    if (counts) counts[i] = (uint32_t)min(c, 0xFFFFFFFFull);
  }
  const uint32_t w = (i & ~63u) >> 6;
  if (zc && w < nw) {
    // a layer's pairs summed over the wave (u32: a wave sum that wraps
    // makes the u64 total exceed 2^32, which the host refuses); lane
    // z - z0 of each 64-layer window keeps its layer's sum
    const int lane = __lane_id();
    const bool any = c > 0;
    const LayerBox bx = layer_box(G, lo, hi);
    const uint32_t per = any ? (uint32_t)(bx.u1 - bx.u0 + 1) * (uint32_t)(bx.v1 - bx.v0 + 1) : 0u;
    int32_t zmin, zmax;
    wave_layer_range(any, bx.l0, bx.l1, &zmin, &zmax);
    const uint32_t nbl = layer_nbl(G);
    for (uint32_t z = (uint32_t)lane; z < nbl; z += 64)
      if ((int32_t)z < zmin || (int32_t)z > zmax) zc[(size_t)z * nw + w] = 0;
    for (int32_t z0 = zmin; z0 <= zmax; z0 += 64) {
      uint32_t mine = 0;
      const int32_t z1 = min(zmax, z0 + 63);
      for (int32_t z = z0; z <= z1; z++) {
        const uint32_t inc = wave_inclusive_scan_dpp((any && bx.l0 <= z && z <= bx.l1) ? per : 0u);
        const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        if (lane == z - z0) mine = v;
      }
      if (z0 + lane <= z1) zc[(size_t)(z0 + lane) * nw + w] = mine;
    }
  }
  // exact u64 total so the host can refuse grids whose u32 offsets would
  // wrap: summed per workgroup and added to one of kTotalSlots slots (every
  // wave adding to one address serialised the adds: 190 us at 1M Gaussians)
  __shared__ unsigned long long s_c[4];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = (s_c[0] + s_c[1]) + (s_c[2] + s_c[3]);
    if (t) atomicAdd(d_total64 + (blockIdx.x % kTotalSlots), t);
  }
}

// The (block, Gaussian) pairs placed layer-major (layer route): the pairs
// of block layer z (along the layer axis) come before those of layer z + 1,
// and inside a layer in Gaussian order, each Gaussian's blocks in (v, u)
// order. A pair's key is ((v nbu + u) << lb) | z. Wave w writes its
// Gaussians' pairs of layer z from zbase[z * nw + w], consecutive lanes on
// consecutive pairs: a lane finds its pair's Gaussian among the few with
// blocks in the layer by reading their run starts with readlane (no LDS
// round trips). A stable sort on the key's bits from lb up then orders the
// pairs by (v, u, z), Gaussian order inside each block.
__global__ __launch_bounds__(256) void k_vox_emit_layers(const VGauss *__restrict__ rec, uint32_t n, VoxGrid G,
                                                         const uint32_t *__restrict__ zbase, uint32_t nw,
                                                         uint32_t lb, uint32_t npairs, uint32_t *__restrict__ keys,
                                                         uint32_t *__restrict__ vals) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t i0 = xcd_chunk() * 256 + (uint32_t)wave * 64u;
  const uint32_t w = i0 >> 6;
  if (w >= nw) return;
  const uint32_t i = i0 + (uint32_t)lane;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  if (i < n) {
    const VGauss g = rec[i];
    axis_block_range(G, 0, g.px, g.ex, &lo[0], &hi[0]);
    axis_block_range(G, 1, g.py, g.ey, &lo[1], &hi[1]);
    axis_block_range(G, 2, g.pz, g.ez, &lo[2], &hi[2]);
  }
  const bool any = i < n && hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2];
  const LayerBox bx = layer_box(G, lo, hi);
  const uint32_t nbu = layer_nbu(G);
  const uint32_t wx = any ? (uint32_t)(bx.u1 - bx.u0 + 1) : 1u;
  const uint32_t per = any ? wx * (uint32_t)(bx.v1 - bx.v0 + 1) : 0u;
  const uint32_t xy0 = any ? (uint32_t)bx.v0 * nbu + (uint32_t)bx.u0 : 0u;
  int32_t zmin, zmax;
  wave_layer_range(any, bx.l0, bx.l1, &zmin, &zmax);
  for (int32_t zw = zmin; zw <= zmax; zw += 64) {
    // the run starts of this 64-layer window, lane z - zw holding layer z's
    const int32_t zl = zw + lane;
    const uint32_t zbl = zl <= zmax ? zbase[(size_t)zl * nw + w] : 0u;
    const int32_t ze = min(zmax, zw + 63);
    for (int32_t z = zw; z <= ze; z++) {
      const uint32_t m = (any && bx.l0 <= z && z <= bx.l1) ? per : 0u;
      const uint64_t cov = __ballot(m > 0u);
      if (cov == 0) continue;
      // exclusive scan of m over the lanes: each lane's run start
      const uint32_t excl = wave_inclusive_scan_dpp(m) - m;
      const uint32_t run = (uint32_t)__builtin_amdgcn_readlane((int)(excl + m), 63);
      const uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)zbl, z - zw);
      for (uint32_t q0 = 0; q0 < run; q0 += 64u) {
        const uint32_t q = q0 + (uint32_t)lane;
        // the last lane with pairs whose run starts at or before q
        uint32_t gl = 0, gstart = 0, gwx = 1, gxy0 = 0;
        uint64_t t = cov;
        while (t) {
          const int l = __builtin_ctzll(t);
          t &= t - 1ull;
          const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)excl, l);
          const uint32_t lwx = (uint32_t)__builtin_amdgcn_readlane((int)wx, l);
          const uint32_t lxy0 = (uint32_t)__builtin_amdgcn_readlane((int)xy0, l);
          if (st <= q) {
            gl = (uint32_t)l;
            gstart = st;
            gwx = lwx;
            gxy0 = lxy0;
          }
        }
        if (q < run && base + q < npairs) {
          const uint32_t local = q - gstart;
          const uint32_t y = local / gwx;
          const uint32_t x = local - y * gwx;
          keys[base + q] = ((gxy0 + y * nbu + x) << lb) | (uint32_t)z;
          vals[base + q] = i0 + gl;
        }
      }
    }
  }
}

// The (block, Gaussian) pairs in Gaussian order, each Gaussian's blocks in
// (z, y, x) order from offs[i]. A wave takes 64 Gaussians, whose pairs are
// one contiguous run of the output, and writes that run with consecutive
// lanes on consecutive pairs: each lane finds its pair's Gaussian by binary
// search over the 64 starts and its block from the index inside the
// Gaussian's box. (A lane writing its own Gaussian's pairs left every store
// instruction 64 scattered runs: 393 us at 1M Gaussians.)
__global__ __launch_bounds__(256) void k_vox_emit(const VGauss *__restrict__ rec, uint32_t n, VoxGrid G,
                                                  const uint32_t *__restrict__ offs, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
  __shared__ uint32_t s_start[4][65];
  __shared__ uint32_t s_x0[4][64], s_y0[4][64], s_z0[4][64], s_wx[4][64], s_wy[4][64];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t i0 = blockIdx.x * 256 + (uint32_t)wave * 64u;
  if (i0 >= n) return;
  const uint32_t i = i0 + (uint32_t)lane;
  uint32_t cnt = 0, wx = 1, wy = 1;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  if (i < n) {
    const VGauss g = rec[i];
    axis_block_range(G, 0, g.px, g.ex, &lo[0], &hi[0]);
    axis_block_range(G, 1, g.py, g.ey, &lo[1], &hi[1]);
    axis_block_range(G, 2, g.pz, g.ez, &lo[2], &hi[2]);
    if (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]) {
      wx = (uint32_t)(hi[0] - lo[0] + 1);
      wy = (uint32_t)(hi[1] - lo[1] + 1);
      cnt = wx * wy * (uint32_t)(hi[2] - lo[2] + 1);
    }
  }
  const uint32_t o0 = offs[i0];
  const uint32_t last = min(n - i0, 64u) - 1u;
  const uint32_t rel = i < n ? offs[i] - o0 : 0u;
  s_start[wave][lane] = rel;
  s_x0[wave][lane] = (uint32_t)lo[0];
  s_y0[wave][lane] = (uint32_t)lo[1];
  s_z0[wave][lane] = (uint32_t)lo[2];
  s_wx[wave][lane] = wx;
  s_wy[wave][lane] = wy;
  // run length: the last Gaussian's start plus its count
  const uint32_t total = __shfl(rel + cnt, (int)last, 64);
  if (lane == 0) s_start[wave][64] = 0xFFFFFFFFu;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const uint32_t nbx = (uint32_t)G.nb[0], nby = (uint32_t)G.nb[1];
  for (uint32_t q = (uint32_t)lane; q < total; q += 64u) {
    // the last Gaussian whose start is <= q (Gaussians without pairs share
    // the next one's start and are passed over)
    uint32_t g = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1)
      if (g + step <= last && s_start[wave][g + step] <= q) g += step;
    const uint32_t local = q - s_start[wave][g];
    const uint32_t gx = s_wx[wave][g], gy = s_wy[wave][g];
    const uint32_t t = local / gx;
    const uint32_t x = local - t * gx;
    const uint32_t z = t / gy;
    const uint32_t y = t - z * gy;
    keys[o0 + q] = (s_x0[wave][g] + x) + (s_y0[wave][g] + y) * nbx + (s_z0[wave][g] + z) * nbx * nby;
    vals[o0 + q] = i0 + g;
  }
}

// evalGaussian (voxel/shaders/voxelize.ts:62-84), reference order of operations
// One Gaussian's contribution at voxel centre (vx, vy, vz): the reference's
// per-voxel body (voxel/shaders/voxelize.ts:62-84), with the record as
// staged in LDS, where ex, ey, ez hold ex + half, ey + half, ez + half (the
// same f32 sums, made once per staged Gaussian), and lo/hi the voxel's
// centre -+ half. Exact rewrites, none of which changes a mask bit:
// - the three box tests are evaluated together;
// - the clamp is one median-of-three instruction (lo < hi, so the median
//   of (p, lo, hi) is min(max(p, lo), hi)): it differs from WGSL's
//   compare-select only in the sign of a zero result when both operands
//   are zero, which the squares of the distance remove, and the bound
//   operands are finite, so a NaN centre gives NaN either way (the
//   distance subtracts the NaN centre);
// - e^x of x = -dd/2 <= 0 (or NaN) is dl_expf_nonpos after clamping x at
//   -104 as dl_expf does, bit-identical to dl_expf there.
__device__ inline float eval_gauss(float vx, float vy, float vz, const float lo[3], const float hi[3],
                                   const VGauss &g) {
  const float d0 = vx - g.px, d1 = vy - g.py, d2 = vz - g.pz;
  const bool out = (__builtin_fabsf(d0) > g.ex) | (__builtin_fabsf(d1) > g.ey) | (__builtin_fabsf(d2) > g.ez);
  if (out) return 0.0f;
  const float c0 = __builtin_amdgcn_fmed3f(g.px, lo[0], hi[0]) - g.px;
  const float c1 = __builtin_amdgcn_fmed3f(g.py, lo[1], hi[1]) - g.py;
  const float c2 = __builtin_amdgcn_fmed3f(g.pz, lo[2], hi[2]) - g.pz;
  const float qx = -g.qx, qy = -g.qy, qz = -g.qz, qw = g.qw;
  const float t0 = 2.0f * (qy * c2 - qz * c1);
  const float t1 = 2.0f * (qz * c0 - qx * c2);
  const float t2 = 2.0f * (qx * c1 - qy * c0);
  const float l0 = (c0 + qw * t0) + (qy * t2 - qz * t1);
  const float l1 = (c1 + qw * t1) + (qz * t0 - qx * t2);
  const float l2 = (c2 + qw * t2) + (qx * t1 - qy * t0);
  const float s0 = l0 * g.is0, s1 = l1 * g.is1, s2 = l2 * g.is2;
  const float dd = (s0 * s0 + s1 * s1) + s2 * s2;
  const float x = (-0.5f) * dd;
  return g.opacity * dl_expf_nonpos(x < -104.0f ? -104.0f : x);
}

// One wave64 per block; lane = linear voxel index z*16 + y*4 + x, so the
// solid ballot is the reference's 64-bit block mask (voxel/shaders/voxelize.ts:137-140).
__global__ __launch_bounds__(64) void k_vox_eval(const VGauss *__restrict__ rec, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ vals, VoxGrid G,
                                                 uint2 *__restrict__ masks) {
  __shared__ VGauss s_g[64];
  const uint32_t blk = blockIdx.x;
  const uint32_t nbx = (uint32_t)G.nb[0], nby = (uint32_t)G.nb[1];
  const uint32_t ax = blk % nbx, ay = (blk / nbx) % nby, az = blk / (nbx * nby);
  // the block's list: its flat index on the flat route, its (v, u, layer)
  // key on the layer route
  const uint32_t key = G.layered ? layer_key(G, ax, ay, az) : blk;
  const uint32_t start = off[key], end = off[key + 1];
  const int lane = threadIdx.x;
  if (start == end) {
    if (lane == 0) masks[blk] = make_uint2(0u, 0u);
    return;
  }
  const uint32_t a3[3] = {ax, ay, az};
  const uint32_t l3[3] = {(uint32_t)lane & 3u, ((uint32_t)lane >> 2) & 3u, (uint32_t)lane >> 4};
  float vc[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint32_t b0 = (a3[k] / 16u) * 16u;                        // batch origin (voxelize.ts:186-195)
    const float bmin = (float)(G.bmn[k] + (double)b0 * G.bs);       // BatchInfo.blockMin, f32
    const float boff = ((float)(a3[k] - b0) * 4.0f) * G.res;        // voxel/shaders/voxelize.ts:104
    vc[k] = (bmin + boff) + (((float)l3[k] + 0.5f) * G.res);       // voxel/shaders/voxelize.ts:105
  }
  float lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = vc[k] - G.half;
    hi[k] = vc[k] + G.half;
  }
  // the sum ends at G.stop: the reference's 7, or lower where the voxel is
  // already solid whatever terms follow (voxel_solid_stop)
  float sigma = 0.0f;
  for (uint32_t base = start; base < end; base += 64) {
    if (__ballot(sigma < G.stop) == 0) break;
    const uint32_t i = base + lane;
    if (i < end) {
      VGauss g = rec[vals[i]];
      g.ex = g.ex + G.half;
      g.ey = g.ey + G.half;
      g.ez = g.ez + G.half;
      s_g[lane] = g;
    }
    __syncthreads();
    const uint32_t cnt = min(64u, end - base);
    if (sigma < G.stop) {
      for (uint32_t c = 0; c < cnt; c++) {
        sigma = sigma + eval_gauss(vc[0], vc[1], vc[2], lo, hi, s_g[c]);
        if (sigma >= G.stop) break;
      }
    }
    __syncthreads();
  }
  const float opacity = 1.0f - dl_expf(-sigma);
  const uint64_t solid = __ballot(opacity >= G.cutoff);
  if (lane == 0) masks[blk] = make_uint2((uint32_t)solid, (uint32_t)(solid >> 32));
}

struct GridDims {
  int32_t nb[3];
};

// Insertion order of BlockMaskBuffer (voxelize.ts:186-188, 162-177): batches
// in (bz, by, bx) loop order, blocks inside a batch by lx + ly*cx + lz*cx*cy.
__device__ inline uint32_t ref_order(uint32_t ax, uint32_t ay, uint32_t az, const GridDims &G) {
  const uint32_t nbx = (uint32_t)G.nb[0], nby = (uint32_t)G.nb[1], nbz = (uint32_t)G.nb[2];
  const uint32_t bx0 = (ax / 16u) * 16u, by0 = (ay / 16u) * 16u, bz0 = (az / 16u) * 16u;
  const uint32_t cx = min(16u, nbx - bx0), cy = min(16u, nby - by0), cz = min(16u, nbz - bz0);
  const uint32_t local = (ax - bx0) + (ay - by0) * cx + (az - bz0) * cx * cy;
  return bz0 * nby * nbx + cz * by0 * nbx + cz * cy * bx0 + local;
}

__global__ __launch_bounds__(256) void k_vox_order(const uint2 *__restrict__ masks, uint32_t total, GridDims G,
                                                   uint32_t *__restrict__ flags, uint2 *__restrict__ omask,
                                                   uint32_t *__restrict__ oidx) {
  const uint32_t blk = blockIdx.x * 256 + threadIdx.x;
  if (blk >= total) return;
  const uint32_t nbx = (uint32_t)G.nb[0], nby = (uint32_t)G.nb[1];
  const uint32_t ax = blk % nbx, ay = (blk / nbx) % nby, az = blk / (nbx * nby);
  const uint32_t o = ref_order(ax, ay, az, G);
  const uint2 m = masks[blk];
  omask[o] = m;
  oidx[o] = blk;
  flags[o] = (m.x | m.y) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_vox_compact(const uint32_t *__restrict__ flags,
                                                     const uint32_t *__restrict__ pos, const uint2 *__restrict__ omask,
                                                     const uint32_t *__restrict__ oidx, uint32_t total,
                                                     uint32_t *__restrict__ out_idx, uint32_t *__restrict__ out_lo,
                                                     uint32_t *__restrict__ out_hi) {
  const uint32_t o = blockIdx.x * 256 + threadIdx.x;
  if (o >= total || !flags[o]) return;
  const uint32_t d = pos[o];
  out_idx[d] = oidx[o];
  out_lo[d] = omask[o].x;
  out_hi[d] = omask[o].y;
}

double js_round(double x) {
  if (!std::isfinite(x)) return x;
  const double f = std::floor(x);
  return (x - f >= 0.5) ? f + 1.0 : f;
}

}  // namespace

VoxelScratch::~VoxelScratch() {
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (host_small) (void)hipHostFree(host_small);
}

void VoxelScratch::ensure_events() {
  for (auto &e : ev)
    if (!e) DL_HIP(hipEventCreate(&e));
  if (!host_small) DL_HIP(hipHostMalloc(&host_small, kHostSmallBytes, hipHostMallocDefault));
}

void voxel_compact(Context &c, hipStream_t s, VoxelResult &r, hipEvent_t done) {
  r.count = 0;
  r.host_ready = false;
  if (r.total == 0) {
    if (done) DL_HIP(hipEventRecord(done, s));
    DL_HIP(hipStreamSynchronize(s));
    return;
  }
  if (r.total >= 0xFFFFFFFFull) fail(DL_ERR_CAPACITY, "voxel grid has more than 2^32-1 blocks");
  const uint32_t total = (uint32_t)r.total;
  GridDims G = {{r.nb[0], r.nb[1], r.nb[2]}};
  DeviceBuffer &flags = c.vox.flags, &fpos = c.vox.fpos, &omask = c.vox.omask, &oidx = c.vox.oidx;
  DeviceBuffer &scan_ws = c.vox.cscan_ws, &small = c.vox.csmall;
  flags.reserve((size_t)total * 4);
  fpos.reserve((size_t)total * 4);
  omask.reserve((size_t)total * 8);
  oidx.reserve((size_t)total * 4);
  scan_ws.reserve(scan_workspace_bytes(total));
  small.reserve(256);
  k_vox_order<<<div_up(total, 256), 256, 0, s>>>(r.masks.as<uint2>(), total, G, flags.as<uint32_t>(),
                                                 omask.as<uint2>(), oidx.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32(s, flags.as<uint32_t>(), fpos.as<uint32_t>(), total, small.as<uint32_t>(), scan_ws.get());
  // the lists are sized for every block, so the compaction follows the scan
  // without a host round trip for the non-empty count
  r.idx.reserve((size_t)total * 4);
  r.lo.reserve((size_t)total * 4);
  r.hi.reserve((size_t)total * 4);
  k_vox_compact<<<div_up(total, 256), 256, 0, s>>>(flags.as<uint32_t>(), fpos.as<uint32_t>(), omask.as<uint2>(),
                                                   oidx.as<uint32_t>(), total, r.idx.as<uint32_t>(),
                                                   r.lo.as<uint32_t>(), r.hi.as<uint32_t>());
  DL_HIP(hipGetLastError());
  if (done) DL_HIP(hipEventRecord(done, s));
  uint32_t nonempty = 0;
  DL_HIP(hipMemcpyAsync(&nonempty, small.get(), 4, hipMemcpyDeviceToHost, s));
  DL_HIP(hipStreamSynchronize(s));
  r.count = nonempty;
}

void voxel_lists_to_host(const VoxelResult &r) {
  if (r.host_ready) return;
  r.h_idx.resize(r.count);
  r.h_lo.resize(r.count);
  r.h_hi.resize(r.count);
  if (r.count) {
    DL_HIP(hipSetDevice(r.device));
    DL_HIP(hipMemcpy(r.h_idx.data(), r.idx.get(), (size_t)r.count * 4, hipMemcpyDeviceToHost));
    DL_HIP(hipMemcpy(r.h_lo.data(), r.lo.get(), (size_t)r.count * 4, hipMemcpyDeviceToHost));
    DL_HIP(hipMemcpy(r.h_hi.data(), r.hi.get(), (size_t)r.count * 4, hipMemcpyDeviceToHost));
  }
  r.host_ready = true;
}

static void voxelize_impl(Context &ctx, const dl_splats &sp, const dl_voxel_opts *opts, VoxelResult &out) {
  const double vres = (opts && opts->voxel_resolution > 0) ? opts->voxel_resolution : 0.05;
  const double cutoff = (opts && opts->opacity_cutoff >= 0) ? opts->opacity_cutoff : 0.1;
  const double bs = 4.0 * vres;
  const uint32_t n = sp.count;
  ctx.make_current();
  out.device = ctx.device;
  out.vres = vres;
  hipStream_t s = ctx.stream;
  VoxelScratch &W = ctx.vox;
  W.ensure_events();
  hipEvent_t e0 = W.ev[0], e1 = W.ev[1];

  DeviceBuffer &pos = W.pos, &rot = W.rot, &ls = W.ls, &op = W.op, &rec = W.rec, &partials = W.partials;
  std::vector<double> part;
  const uint32_t nprep = std::max<uint32_t>(1, div_up(n, kPrepBlock));
  if (n > 0) {
    if (!sp.positions || !sp.rotations || !sp.log_scales || !sp.opacity_logits)
      fail(DL_ERR_INVALID, "null splat array");
    pos.reserve((size_t)n * 12);
    rot.reserve((size_t)n * 16);
    ls.reserve((size_t)n * 12);
    op.reserve((size_t)n * 4);
    rec.reserve((size_t)n * sizeof(VGauss));
    partials.reserve(((size_t)nprep + 1) * 6 * sizeof(double));
    ctx.staging.to_device(pos.get(), sp.positions, (size_t)n * 12, s);
    ctx.staging.to_device(rot.get(), sp.rotations, (size_t)n * 16, s);
    ctx.staging.to_device(ls.get(), sp.log_scales, (size_t)n * 12, s);
    ctx.staging.to_device(op.get(), sp.opacity_logits, (size_t)n * 4, s);
    DL_HIP(hipEventRecord(e0, s));
    k_vox_prep<<<nprep, kPrepBlock, 0, s>>>(pos.as<float>(), rot.as<float>(), ls.as<float>(), op.as<float>(), n,
                                           rec.as<VGauss>(), partials.as<double>());
    DL_HIP(hipGetLastError());
    k_vox_bounds<<<1, 1024, 0, s>>>(partials.as<double>(), nprep, partials.as<double>() + (size_t)nprep * 6);
    DL_HIP(hipGetLastError());
    // the pair total's slots are cleared before the host waits for the
    // bounds, not after (one enqueue less between the wait and the count)
    W.small.reserve(64 + kTotalSlots * sizeof(unsigned long long));
    DL_HIP(hipMemsetAsync(W.small.as<uint8_t>() + 64, 0, kTotalSlots * sizeof(unsigned long long), s));
    part.resize(6);
    DL_HIP(hipMemcpyAsync(W.host_small, partials.as<double>() + (size_t)nprep * 6, 6 * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    std::memcpy(part.data(), W.host_small, 6 * sizeof(double));
  } else {
    DL_HIP(hipEventRecord(e0, s));
  }

  double bmn[3], bmx[3];
  if (opts && opts->has_bounds) {
    for (int k = 0; k < 3; k++) {
      bmn[k] = opts->bounds_min[k];
      bmx[k] = opts->bounds_max[k];
    }
  } else {
    // computeBlockAlignedBounds (voxelize.ts:278-306)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (n)
      for (int k = 0; k < 3; k++) {
        mn[k] = part[k];
        mx[k] = part[3 + k];
      }
    for (int k = 0; k < 3; k++) {
      bmn[k] = std::floor(mn[k] / bs) * bs;
      bmx[k] = std::ceil(mx[k] / bs) * bs;
    }
  }
  double nbd[3];
  for (int k = 0; k < 3; k++) {
    nbd[k] = js_round((bmx[k] - bmn[k]) / bs);
    out.bmin[k] = bmn[k];
    out.bmax[k] = bmx[k];
    out.nblocks[k] = nbd[k];
  }
  out.count = 0;
  out.total = 0;
  out.host_ready = false;
  int64_t nb[3];
  for (int k = 0; k < 3; k++) nb[k] = (std::isfinite(nbd[k]) && nbd[k] > 0) ? (int64_t)nbd[k] : 0;
  const uint64_t total64 = (uint64_t)nb[0] * (uint64_t)nb[1] * (uint64_t)nb[2];
  if (n == 0 || total64 == 0) {
    DL_HIP(hipEventRecord(e1, s));
    DL_HIP(hipStreamSynchronize(s));
    DL_HIP(hipEventElapsedTime(&out.device_ms, e0, e1));
    return;
  }
  if (total64 >= 0xFFFFFFFFull || nb[0] > 0x7FFFFFFF || nb[1] > 0x7FFFFFFF || nb[2] > 0x7FFFFFFF)
    fail(DL_ERR_CAPACITY, "voxel grid has more than 2^32-1 blocks");
  const uint32_t total = (uint32_t)total64;
  for (int k = 0; k < 3; k++) out.nb[k] = (int32_t)nb[k];
  out.total = total;

  VoxGrid G;
  for (int k = 0; k < 3; k++) {
    G.bmn[k] = bmn[k];
    G.nb[k] = (int32_t)nb[k];
    G.nbatch[k] = (uint32_t)((nb[k] + 15) / 16);
  }
  G.bs = bs;
  G.res = (float)vres;
  G.half = G.res * 0.5f;
  G.cutoff = (float)cutoff;
  G.stop = voxel_solid_stop(G.cutoff);

  // Two routes to the per-block lists (each block's Gaussians in index
  // order). Layer route: the layers are the blocks along the axis of fewest
  // blocks (l; u and v the other two axes in order); the pairs are placed
  // layer by layer and keyed ((v nbu + u) << lb) | l, so a stable sort on
  // the key's bits from lb up (two passes on the workload's 14 bits,
  // against three on the 20-bit flat index) orders them by block. Its
  // counting and placement cost grows with the layers a wave's Gaussians
  // span (each wave loops over them), so it is taken for grids of at most
  // 64 layers. Flat route otherwise: pairs in Gaussian order keyed by the
  // flat block index, sorted on every bit.
  const uint32_t nw = div_up(n, 64);
  uint32_t lax = 0;
  for (uint32_t k = 1; k < 3; k++)
    if (nb[k] < nb[lax]) lax = k;
  const uint32_t nbl = (uint32_t)nb[lax];
  const uint32_t nbu = (uint32_t)nb[lax == 0 ? 1 : 0], nbv = (uint32_t)nb[lax == 2 ? 1 : 2];
  const uint32_t lb = bit_width_u32(nbl - 1);
  const uint32_t uvbits = bit_width_u32(nbu * nbv - 1);
  const bool layered = nbl <= 64 && lb + uvbits <= 31 && (uint64_t)nbl * nw <= (1ull << 26);
  G.layered = layered ? 1u : 0u;
  G.lax = lax;
  G.lb = layered ? lb : 0u;
  DeviceBuffer &counts = W.counts, &offs = W.offs, &scan_ws = W.scan_ws, &small = W.small;
  const size_t nzc = (size_t)nbl * nw;
  if (layered) {
    counts.reserve(nzc * 4);
    offs.reserve(nzc * 4);
    scan_ws.reserve(scan_workspace_bytes(nzc));
  } else {
    counts.reserve((size_t)n * 4);
    offs.reserve((size_t)n * 4);
    scan_ws.reserve(scan_workspace_bytes(n));
  }
  // small: [0] the scan's u32 total; from byte 64 the u64 total's slots
  // (cleared above)
  unsigned long long *d_total64 = reinterpret_cast<unsigned long long *>(small.as<uint8_t>() + 64);
  // the layer route deals contiguous runs of chunks to each XCD (xcd_chunk)
  const uint32_t vgrid = layered ? 8u * div_up(div_up(n, 256), 8) : div_up(n, 256);
  k_vox_count<<<vgrid, 256, 0, s>>>(rec.as<VGauss>(), n, G, layered ? nullptr : counts.as<uint32_t>(), d_total64,
                                    layered ? counts.as<uint32_t>() : nullptr, nw);
  DL_HIP(hipGetLastError());
  // the u64 total and the scan's u32 total in one read: a total above 2^32
  // fails before the (wrapped) scan is used
  exclusive_scan_u32(s, counts.as<uint32_t>(), offs.as<uint32_t>(), layered ? nzc : n, small.as<uint32_t>(),
                     scan_ws.get());
  uint32_t head[16 + 2 * kTotalSlots];
  static_assert(sizeof(head) <= VoxelScratch::kHostSmallBytes, "pinned readback too small");
  DL_HIP(hipMemcpyAsync(W.host_small, small.get(), sizeof(head), hipMemcpyDeviceToHost, s));
  DL_HIP(hipStreamSynchronize(s));
  std::memcpy(head, W.host_small, sizeof(head));
  unsigned long long total_pairs = 0;
  for (uint32_t k = 0; k < kTotalSlots; k++)
    total_pairs += (unsigned long long)head[16 + 2 * k] | ((unsigned long long)head[17 + 2 * k] << 32);
  if (total_pairs > 0xFFFFFFF0ull) fail(DL_ERR_CAPACITY, "more than 2^32 (block, Gaussian) pairs");
  const uint32_t npairs = head[0];
  // the layer counts' total (wave scans) must equal the per-Gaussian total
  if (npairs != total_pairs) fail(DL_ERR_INTERNAL, "voxelizer: the layer counts disagree with the pair total");

  DeviceBuffer &keys = W.keys, &vals = W.vals, &tkeys = W.tkeys, &tvals = W.tvals, &sort_ws = W.sort_ws;
  DeviceBuffer &ranges = W.ranges;
  DeviceBuffer &masks = out.masks;
  // list offsets per block key: total blocks, or the (v, u, l) key space
  const uint32_t nbins = layered ? ((nbu * nbv - 1u) << lb) + (1u << lb) : total;
  const uint32_t key_bits = layered ? lb + uvbits : std::max<uint32_t>(1, bit_width_u32(total - 1));
  keys.reserve((size_t)std::max<uint32_t>(npairs, 1) * 4);
  vals.reserve((size_t)std::max<uint32_t>(npairs, 1) * 4);
  tkeys.reserve((size_t)std::max<uint32_t>(npairs, 1) * 4);
  tvals.reserve((size_t)std::max<uint32_t>(npairs, 1) * 4);
  sort_ws.reserve(radix_workspace_bytes(std::max<uint32_t>(npairs, 1), (int)key_bits));
  ranges.reserve(((size_t)nbins + 1) * 4);
  const uint32_t *skeys = keys.as<uint32_t>(), *svals = vals.as<uint32_t>();
  if (npairs > 0) {
    if (layered)
      k_vox_emit_layers<<<vgrid, 256, 0, s>>>(rec.as<VGauss>(), n, G, offs.as<uint32_t>(), nw, lb, npairs,
                                              keys.as<uint32_t>(), vals.as<uint32_t>());
    else
      k_vox_emit<<<div_up(n, 256), 256, 0, s>>>(rec.as<VGauss>(), n, G, offs.as<uint32_t>(), keys.as<uint32_t>(),
                                                vals.as<uint32_t>());
    DL_HIP(hipGetLastError());
    const bool in_tmp = radix_sort_pairs(s, keys.as<uint32_t>(), vals.as<uint32_t>(), tkeys.as<uint32_t>(),
                                         tvals.as<uint32_t>(), npairs, (int)key_bits, sort_ws.get(), 9,
                                         RadixKeyMap{}, false, layered ? (int)lb : 0);
    if (in_tmp) {
      skeys = tkeys.as<uint32_t>();
      svals = tvals.as<uint32_t>();
    }
  }
  bin_ranges(s, skeys, npairs, nbins, ranges.as<uint32_t>());
  // the result's masks are allocated (a new result per call) once the pair
  // build and sort are queued, so the allocation runs while they do
  masks.reserve((size_t)total * 8);
  hipEvent_t ev_a = W.ev[2], ev_b = W.ev[3];
  DL_HIP(hipEventRecord(ev_a, s));
  k_vox_eval<<<total, 64, 0, s>>>(rec.as<VGauss>(), ranges.as<uint32_t>(), svals, G, masks.as<uint2>());
  DL_HIP(hipGetLastError());
  DL_HIP(hipEventRecord(ev_b, s));
  out.block_pairs = npairs;

  // reference insertion order + compaction of non-empty blocks
  voxel_compact(ctx, s, out, e1);
  DL_HIP(hipEventElapsedTime(&out.device_ms, e0, e1));
  DL_HIP(hipEventElapsedTime(&out.eval_ms, ev_a, ev_b));
}

}  // namespace dl

using namespace dl;

extern "C" {

int dl_voxelize(dl_context *ctx, const dl_splats *splats, const dl_voxel_opts *opts, dl_voxel_result **out) {
  return guard([&] {
    if (!ctx || !splats || !out) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<dl_voxel_result> r(new dl_voxel_result());
    voxelize_impl(ctx->c, *splats, opts, r->r);
    *out = r.release();
  });
}

uint64_t dl_voxel_result_count(const dl_voxel_result *res) { return res ? res->r.count : 0; }

int dl_voxel_result_blocks(const dl_voxel_result *res, uint32_t *block_idx, uint32_t *mask_lo, uint32_t *mask_hi) {
  return guard([&] {
    if (!res || (res->r.count && (!block_idx || !mask_lo || !mask_hi))) fail(DL_ERR_INVALID, "null argument");
    const size_t n = res->r.count;
    if (!n) return;
    voxel_lists_to_host(res->r);
    std::memcpy(block_idx, res->r.h_idx.data(), n * 4);
    std::memcpy(mask_lo, res->r.h_lo.data(), n * 4);
    std::memcpy(mask_hi, res->r.h_hi.data(), n * 4);
  });
}

int dl_voxel_result_bounds(const dl_voxel_result *res, double bounds_min[3], double bounds_max[3],
                           double num_blocks[3]) {
  return guard([&] {
    if (!res || !bounds_min || !bounds_max || !num_blocks) fail(DL_ERR_INVALID, "null argument");
    for (int k = 0; k < 3; k++) {
      bounds_min[k] = res->r.bmin[k];
      bounds_max[k] = res->r.bmax[k];
      num_blocks[k] = res->r.nblocks[k];
    }
  });
}

float dl_voxel_result_device_ms(const dl_voxel_result *res) { return res ? res->r.device_ms : 0.0f; }

int dl_voxel_result_stats(const dl_voxel_result *res, dl_voxel_stats *out) {
  return guard([&] {
    if (!res || !out) fail(DL_ERR_INVALID, "null argument");
    *out = dl_voxel_stats{};
    out->device_ms = res->r.device_ms;
    out->eval_ms = res->r.eval_ms;
    out->block_pairs = res->r.block_pairs;
    out->grid_blocks = res->r.total;
  });
}

void dl_voxel_result_free(dl_voxel_result *res) { delete res; }

}  // extern "C"
