// Collision stages over a voxel grid resident in HBM: fill, carve, point
// queries and the quad mesh, plus the upload of a host BlockMaskBuffer.
//
// The reference has no code for these stages. README.md:108 and
// src/voxel/block-mask-buffer.ts:10-11 name fill, carve and mesh as grid
// arithmetic downstream of the voxelizer's BlockMaskBuffer, so their
// definitions are this library's:
//
//   flood  R = the empty voxels joined through 6-connected empty voxels to a
//          seed; the output grid is the complement of R.
//   fill   seeds = every empty voxel on the grid's boundary, so every cavity
//          that solid voxels enclose is closed.
//   carve  seeds = the voxels holding the given world points
//          (floor((p - bmin) / res) in f64; points outside the grid or in
//          solid voxels seed nothing), so every voxel that cannot be reached
//          from the points becomes solid.
//   mesh   a face of direction d (+x, -x, +y, -y, +z, -z) is a solid voxel
//          whose neighbour along d is empty or outside the grid. With normal
//          axis a and in-plane axes (u, v) = (y, z), (x, z), (x, y) for
//          a = x, y, z, the faces of a slice and row v form maximal runs
//          [u0, u1) along u; a run starts a quad unless row v-1 holds exactly
//          the same run, and the quad covers every following row that holds
//          exactly that run. Quads are ordered by (d, slice, v0, u0); each has
//          the vertices (u0,v0) (u1,v0) (u1,v1) (u0,v1) on the plane
//          slice + (d positive), coordinates (float)(bmin + (double)i * res),
//          and two triangles wound counter-clockwise seen from the side d
//          points to.
//
// Device design.
//   Flood: the dense grid holds one 64-bit mask per 4x4x4 block (voxel bit
//   x + 4y + 16z). A 256-thread workgroup owns an 8x8x4-block tile and keeps
//   it with a one-block halo in LDS; each thread grows its block's reached
//   set with shifts and masks (inside the block) and the neighbours' facing
//   bit planes (across blocks) until the whole tile stops changing, then
//   writes back what changed. Launches repeat until one changes nothing;
//   a tile runs only when it or a face neighbour changed since the launch
//   before, so after the first launch only the moving front costs work.
//   Reached sets only grow, so reading a neighbour's old or new word is
//   always a subset of the answer and the fixed point is unique.
//   Mesh: face bits are gathered into rows of 64-bit words along u (one
//   word per thread), then one thread per row walks its runs with
//   count-trailing-zeros, counts the quads it starts, and after a scan of
//   the counts writes them in order.
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

// bits of a block mask on its faces (bit = x + 4y + 16z)
constexpr uint64_t kX0 = 0x1111111111111111ull, kX3 = 0x8888888888888888ull;
constexpr uint64_t kY0 = 0x000F000F000F000Full, kY3 = 0xF000F000F000F000ull;
constexpr uint64_t kZ0 = 0x000000000000FFFFull, kZ3 = 0xFFFF000000000000ull;

constexpr int kTX = 8, kTY = 8, kTZ = 4;                      // tile, in blocks
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;    // with halo
constexpr int kHalo = kHX * kHY * kHZ;
constexpr int kFloodThreads = kTX * kTY * kTZ;
static_assert(kFloodThreads == 256, "one thread per block of the tile");

struct GridGeom {
  int32_t nb[3];
  double bmin[3];
  double res;
};

__device__ inline uint64_t as_u64(uint2 w) { return (uint64_t)w.x | ((uint64_t)w.y << 32); }
__device__ inline uint2 as_uint2(uint64_t v) { return make_uint2((uint32_t)v, (uint32_t)(v >> 32)); }

// neighbours inside the block
__device__ inline uint64_t grow_in_block(uint64_t r) {
  return ((r << 1) & ~kX0) | ((r >> 1) & ~kX3) | ((r << 4) & ~kY0) | ((r >> 4) & ~kY3) | (r << 16) | (r >> 16);
}

__global__ __launch_bounds__(256) void k_fill_u64(uint64_t *__restrict__ p, uint64_t n, uint64_t value) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ __launch_bounds__(64) void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t value) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) p[i] = value;
}

// BlockMaskBuffer.addBlock (block-mask-buffer.ts:30-41) for a whole list:
// repeated blocks OR-merge; indices were checked on the host.
__global__ __launch_bounds__(256) void k_scatter_blocks(const uint32_t *__restrict__ idx,
                                                        const uint32_t *__restrict__ lo,
                                                        const uint32_t *__restrict__ hi, uint64_t n,
                                                        unsigned long long *__restrict__ grid) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long m = (unsigned long long)lo[i] | ((unsigned long long)hi[i] << 32);
  if (m) atomicOr(&grid[idx[i]], m);
}

// voxel of a world point; false outside the grid (and for NaN)
__device__ inline bool point_voxel(const double *p, const GridGeom &g, int64_t v[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double f = floor((p[k] - g.bmin[k]) / g.res);
    if (!(f >= 0.0 && f < (double)(4 * (int64_t)g.nb[k]))) return false;
    v[k] = (int64_t)f;
  }
  return true;
}

__device__ inline uint64_t block_of(const int64_t v[3], const GridGeom &g) {
  return (uint64_t)(v[0] >> 2) + (uint64_t)(v[1] >> 2) * (uint64_t)g.nb[0] +
         (uint64_t)(v[2] >> 2) * (uint64_t)g.nb[0] * (uint64_t)g.nb[1];
}

__device__ inline uint32_t bit_of(const int64_t v[3]) {
  return (uint32_t)((v[0] & 3) + 4 * (v[1] & 3) + 16 * (v[2] & 3));
}

__global__ __launch_bounds__(256) void k_seed(const double *__restrict__ pts, uint64_t n, GridGeom g,
                                              const uint2 *__restrict__ solid,
                                              unsigned long long *__restrict__ reach) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int64_t v[3];
  if (!point_voxel(pts + 3 * i, g, v)) return;
  const uint64_t blk = block_of(v, g);
  const uint64_t bit = 1ull << bit_of(v);
  if (!(as_u64(solid[blk]) & bit)) atomicOr(&reach[blk], (unsigned long long)bit);
}

__global__ __launch_bounds__(256) void k_query(const double *__restrict__ pts, uint64_t n, GridGeom g,
                                               const uint2 *__restrict__ masks, uint8_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int64_t v[3];
  uint8_t r = 0;
  if (point_voxel(pts + 3 * i, g, v)) r = (uint8_t)((as_u64(masks[block_of(v, g)]) >> bit_of(v)) & 1u);
  out[i] = r;
}

__global__ __launch_bounds__(256) void k_complement(const uint2 *__restrict__ in, uint2 *__restrict__ out,
                                                    uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const uint2 w = in[i];
    out[i] = make_uint2(~w.x, ~w.y);
  }
}

// One flood step over an 8x8x4-block tile (launch number `launch`, from 1);
// sets *changed and last_change[tile] = launch when the tile's reached set
// grew. A tile runs only when it or a face neighbour changed in the previous
// launch or later (every tile in launch 1): otherwise its inputs are what it
// already saturated against. `outside` is the reached set of blocks beyond
// the grid.
__global__ __launch_bounds__(256) void k_flood(const uint2 *__restrict__ solid, uint2 *__restrict__ reach,
                                               GridGeom g, uint32_t tiles_x, uint32_t tiles_y, uint32_t tiles_z,
                                               uint64_t outside, uint32_t launch,
                                               uint32_t *__restrict__ last_change, uint32_t *__restrict__ changed) {
  __shared__ uint64_t R[kHalo];
  __shared__ uint32_t s_active;
  const int32_t nbx = g.nb[0], nby = g.nb[1], nbz = g.nb[2];
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
  if (threadIdx.x == 0) {
    // one thread decides, so that the whole workgroup returns or none does
    const uint32_t need = launch - 1u;
    const uint32_t sxy = tiles_x * tiles_y;
    bool act = launch <= 1u || last_change[tile] >= need;
    act = act || (tx > 0 && last_change[tile - 1] >= need) || (tx + 1 < tiles_x && last_change[tile + 1] >= need);
    act = act || (ty > 0 && last_change[tile - tiles_x] >= need) ||
          (ty + 1 < tiles_y && last_change[tile + tiles_x] >= need);
    act = act || (tz > 0 && last_change[tile - sxy] >= need) || (tz + 1 < tiles_z && last_change[tile + sxy] >= need);
    s_active = act ? 1u : 0u;
  }
  __syncthreads();
  if (!s_active) return;
  const int32_t ox = (int32_t)tx * kTX - 1;
  const int32_t oy = (int32_t)ty * kTY - 1;
  const int32_t oz = (int32_t)tz * kTZ - 1;
  for (int i = threadIdx.x; i < kHalo; i += kFloodThreads) {
    const int32_t bx = ox + i % kHX, by = oy + (i / kHX) % kHY, bz = oz + i / (kHX * kHY);
    uint64_t w = outside;
    if (bx >= 0 && by >= 0 && bz >= 0 && bx < nbx && by < nby && bz < nbz)
      w = as_u64(reach[(uint64_t)bx + (uint64_t)by * (uint64_t)nbx + (uint64_t)bz * (uint64_t)nbx * (uint64_t)nby]);
    R[i] = w;
  }
  const int t = threadIdx.x;
  const int32_t lx = t % kTX, ly = (t / kTX) % kTY, lz = t / (kTX * kTY);
  const int32_t bx = ox + 1 + lx, by = oy + 1 + ly, bz = oz + 1 + lz;
  const bool mine = bx < nbx && by < nby && bz < nbz;
  const uint64_t blk = mine ? (uint64_t)bx + (uint64_t)by * (uint64_t)nbx + (uint64_t)bz * (uint64_t)nbx * (uint64_t)nby
                            : 0;
  const uint64_t empty = mine ? ~as_u64(solid[blk]) : 0ull;
  const int h = (lx + 1) + kHX * ((ly + 1) + kHY * (lz + 1));
  __syncthreads();
  const uint64_t r0 = R[h];
  uint64_t r = r0;
  // every pass that continues has grown some block by at least one bit
  for (int pass = 0; pass <= 64 * kFloodThreads; pass++) {
    bool grew = false;
    if (mine) {
      const uint64_t in = ((R[h - 1] & kX3) >> 3) | ((R[h + 1] & kX0) << 3) | ((R[h - kHX] & kY3) >> 12) |
                          ((R[h + kHX] & kY0) << 12) | ((R[h - kHX * kHY] & kZ3) >> 48) |
                          ((R[h + kHX * kHY] & kZ0) << 48);
      uint64_t n = r | (in & empty);
      for (int it = 0; it < 64; it++) {
        const uint64_t m = n | (grow_in_block(n) & empty);
        if (m == n) break;
        n = m;
      }
      if (n != r) {
        r = n;
        R[h] = n;
        grew = true;
      }
    }
    if (!__syncthreads_or(grew)) break;
  }
  const bool any = __syncthreads_or(mine && r != r0);
  if (mine && r != r0) reach[blk] = as_uint2(r);
  if (any && threadIdx.x == 0) {
    last_change[tile] = launch;
    *changed = 1u;
  }
}

// ---------------------------------------------------------------- mesh
constexpr int kUAxis[3] = {1, 0, 0}, kVAxis[3] = {2, 2, 1};

struct MeshDims {
  int32_t V[3];                  // voxels per axis
  uint32_t W[3];                 // words per row for normal axis a
  unsigned long long rowbase[7];  // first row of direction d
  unsigned long long wordbase[7]; // first word of direction d
};

__device__ inline int dir_of(unsigned long long i, const unsigned long long *base) {
  int d = 0;
#pragma unroll
  for (int k = 1; k < 6; k++) d += (i >= base[k]) ? 1 : 0;
  return d;
}

// 4 bits along axis `ua` of a block mask, starting at bit `base`
__device__ inline uint32_t gather4(uint64_t w, uint32_t base, uint32_t stride) {
  return (uint32_t)(((w >> base) & 1u) | (((w >> (base + stride)) & 1u) << 1) |
                    (((w >> (base + 2 * stride)) & 1u) << 2) | (((w >> (base + 3 * stride)) & 1u) << 3));
}

__global__ __launch_bounds__(256) void k_face_words(const uint2 *__restrict__ masks, GridGeom g, MeshDims m,
                                                    unsigned long long *__restrict__ words) {
  const unsigned long long gi = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  if (gi >= m.wordbase[6]) return;
  const int d = dir_of(gi, m.wordbase);
  const int a = d >> 1, ua = kUAxis[a], va = kVAxis[a];
  const int sg = (d & 1) ? -1 : 1;
  const unsigned long long li = gi - m.wordbase[d];
  const uint32_t W = m.W[a];
  const unsigned long long row = li / W;
  const uint32_t w = (uint32_t)(li % W);
  const int32_t v = (int32_t)(row % (unsigned long long)m.V[va]);
  const int32_t s = (int32_t)(row / (unsigned long long)m.V[va]);
  const uint32_t stride = ua == 0 ? 1u : (ua == 1 ? 4u : 16u);
  uint64_t out = 0;
  for (int j = 0; j < 16; j++) {
    const int32_t u = (int32_t)(w * 64) + 4 * j;
    if (u >= m.V[ua]) break;
    int64_t c[3];
    c[a] = s;
    c[ua] = u;
    c[va] = v;
    const uint32_t base = (uint32_t)((c[0] & 3) + 4 * (c[1] & 3) + 16 * (c[2] & 3));
    const uint32_t s4 = gather4(as_u64(masks[block_of(c, g)]), base, stride);
    if (!s4) continue;
    c[a] += sg;
    uint32_t n4 = 0;
    if (c[a] >= 0 && c[a] < m.V[a]) {
      const uint32_t nbase = (uint32_t)((c[0] & 3) + 4 * (c[1] & 3) + 16 * (c[2] & 3));
      n4 = gather4(as_u64(masks[block_of(c, g)]), nbase, stride);
    }
    out |= (uint64_t)(s4 & ~n4) << (4 * j);
  }
  words[gi] = out;
}

__device__ inline bool bit_at(const unsigned long long *row, uint32_t u) { return (row[u >> 6] >> (u & 63)) & 1ull; }

// first set (want = true) or clear bit at position >= u, or n; bits of the
// last word at positions >= n are zero
template <bool WANT_SET>
__device__ inline uint32_t next_bit(const unsigned long long *row, uint32_t u, uint32_t n, uint32_t W) {
  if (u >= n) return n;
  uint32_t w = u >> 6;
  unsigned long long x = (WANT_SET ? row[w] : ~row[w]) & (~0ull << (u & 63));
  while (x == 0) {
    if (++w >= W) return n;
    x = WANT_SET ? row[w] : ~row[w];
  }
  const uint32_t p = (w << 6) + (uint32_t)__builtin_ctzll(x);
  return p < n ? p : n;
}

__device__ inline bool exact_run(const unsigned long long *row, uint32_t u0, uint32_t u1, uint32_t n, uint32_t W) {
  return bit_at(row, u0) && next_bit<false>(row, u0, n, W) == u1 && (u0 == 0 || !bit_at(row, u0 - 1));
}

template <bool EMIT>
__global__ __launch_bounds__(256) void k_mesh_rows(const unsigned long long *__restrict__ words, GridGeom g,
                                                   MeshDims m, uint32_t *__restrict__ counts,
                                                   unsigned long long *__restrict__ total,
                                                   const uint32_t *__restrict__ qoff, float *__restrict__ pos,
                                                   uint32_t *__restrict__ idx) {
  const unsigned long long gi = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  unsigned long long mine = 0;
  if (gi < m.rowbase[6]) {
    const int d = dir_of(gi, m.rowbase);
    const int a = d >> 1, ua = kUAxis[a], va = kVAxis[a];
    const bool positive = !(d & 1);
    const bool ccw = (a == 1) ? !positive : positive;
    const unsigned long long r = gi - m.rowbase[d];
    const uint32_t W = m.W[a];
    const uint32_t n = (uint32_t)m.V[ua];
    const int32_t Vv = m.V[va];
    const int32_t v = (int32_t)(r % (unsigned long long)Vv);
    const int32_t s = (int32_t)(r / (unsigned long long)Vv);
    const unsigned long long *row = words + m.wordbase[d] + r * W;
    uint32_t q = EMIT ? qoff[gi] : 0u;
    uint32_t cnt = 0;
    uint32_t u = 0;
    for (;;) {
      const uint32_t u0 = next_bit<true>(row, u, n, W);
      if (u0 >= n) break;
      const uint32_t u1 = next_bit<false>(row, u0, n, W);
      u = u1;
      if (v > 0 && exact_run(row - W, u0, u1, n, W)) continue;
      cnt++;
      if (EMIT) {
        int32_t v1 = v + 1;
        while (v1 < Vv && exact_run(row + (size_t)(v1 - v) * W, u0, u1, n, W)) v1++;
        const int32_t cu[4] = {(int32_t)u0, (int32_t)u1, (int32_t)u1, (int32_t)u0};
        const int32_t cv[4] = {v, v, v1, v1};
        float *p = pos + (size_t)12 * q;
        for (int k = 0; k < 4; k++) {
          int64_t c[3];
          c[a] = s + (positive ? 1 : 0);
          c[ua] = cu[k];
          c[va] = cv[k];
          for (int j = 0; j < 3; j++) p[3 * k + j] = (float)(g.bmin[j] + (double)c[j] * g.res);
        }
        const uint32_t b = 4u * q;
        uint32_t *t = idx + (size_t)6 * q;
        t[0] = b;
        t[1] = ccw ? b + 1 : b + 2;
        t[2] = ccw ? b + 2 : b + 1;
        t[3] = b;
        t[4] = ccw ? b + 2 : b + 3;
        t[5] = ccw ? b + 3 : b + 2;
        q++;
      }
    }
    if (!EMIT) counts[gi] = cnt;
    mine = cnt;
  }
  if (!EMIT) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mine += __shfl_xor(mine, o, 64);
    if (__lane_id() == 0 && mine) atomicAdd(total, mine);
  }
}

GridGeom geom_of(const VoxelResult &r) {
  GridGeom g;
  for (int k = 0; k < 3; k++) {
    g.nb[k] = r.nb[k];
    g.bmin[k] = r.bmin[k];
  }
  g.res = r.vres;
  return g;
}

void copy_geometry(const VoxelResult &in, VoxelResult &out) {
  for (int k = 0; k < 3; k++) {
    out.bmin[k] = in.bmin[k];
    out.bmax[k] = in.bmax[k];
    out.nblocks[k] = in.nblocks[k];
    out.nb[k] = in.nb[k];
  }
  out.vres = in.vres;
  out.total = in.total;
  out.device = in.device;
}

// fill (outside = all reached, no points) or carve (outside = none, points)
void flood_impl(Context &ctx, const VoxelResult &in, bool fill, const double *points, uint64_t npoints,
                VoxelResult &out) {
  ctx.make_current();
  copy_geometry(in, out);
  out.device = ctx.device;
  hipStream_t s = ctx.stream;
  hipEvent_t e0, e1;
  DL_HIP(hipEventCreate(&e0));
  DL_HIP(hipEventCreate(&e1));
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } events{e0, e1};
  DL_HIP(hipEventRecord(e0, s));
  if (in.total == 0) {
    voxel_compact(ctx, s, out, e1);
    DL_HIP(hipEventElapsedTime(&out.device_ms, e0, e1));
    return;
  }
  const uint64_t total = in.total;
  const GridGeom g = geom_of(in);
  DeviceBuffer reach, pts, flags;
  reach.reserve(total * 8);
  const uint32_t fill_blocks = (uint32_t)((total + 255) / 256);
  k_fill_u64<<<fill_blocks, 256, 0, s>>>(reach.as<uint64_t>(), total, 0ull);
  DL_HIP(hipGetLastError());
  if (npoints) {
    pts.reserve(npoints * 24);
    DL_HIP(hipMemcpyAsync(pts.get(), points, npoints * 24, hipMemcpyHostToDevice, s));
    k_seed<<<(uint32_t)((npoints + 255) / 256), 256, 0, s>>>(pts.as<double>(), npoints, g, in.masks.as<uint2>(),
                                                             reach.as<unsigned long long>());
    DL_HIP(hipGetLastError());
  }
  const uint32_t tx = (uint32_t)div_up((uint32_t)g.nb[0], kTX), ty = (uint32_t)div_up((uint32_t)g.nb[1], kTY);
  const uint32_t tz = (uint32_t)div_up((uint32_t)g.nb[2], kTZ);
  const uint64_t tiles = (uint64_t)tx * ty * tz;
  if (tiles > 0x7FFFFFFFull) fail(DL_ERR_CAPACITY, "voxel grid too large for the flood");
  // launches between reads of the flags: 8, 16, 32, then 64 (a launch that
  // finds no active tile costs a few microseconds)
  constexpr uint32_t kMaxBatch = 64;
  constexpr uint64_t kMaxLaunches = 1000000;
  DeviceBuffer last_change;
  flags.reserve(kMaxBatch * 4);
  last_change.reserve(tiles * 4);
  k_fill_u32<<<(uint32_t)((tiles + 63) / 64), 64, 0, s>>>(last_change.as<uint32_t>(), (uint32_t)tiles, 0u);
  DL_HIP(hipGetLastError());
  const uint64_t outside = fill ? ~0ull : 0ull;
  uint32_t h_flags[kMaxBatch];
  uint32_t launch = 0, batch = 8;
  for (;; batch = std::min(2 * batch, kMaxBatch)) {
    if (launch >= kMaxLaunches) fail(DL_ERR_INTERNAL, "flood did not converge");
    k_fill_u32<<<1, 64, 0, s>>>(flags.as<uint32_t>(), batch, 0u);
    DL_HIP(hipGetLastError());
    for (uint32_t i = 0; i < batch; i++) {
      k_flood<<<(uint32_t)tiles, kFloodThreads, 0, s>>>(in.masks.as<uint2>(), reach.as<uint2>(), g, tx, ty, tz,
                                                        outside, ++launch, last_change.as<uint32_t>(),
                                                        flags.as<uint32_t>() + i);
      DL_HIP(hipGetLastError());
    }
    DL_HIP(hipMemcpyAsync(h_flags, flags.get(), (size_t)batch * 4, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    if (!h_flags[batch - 1]) break;
  }
  out.masks.reserve(total * 8);
  k_complement<<<fill_blocks, 256, 0, s>>>(reach.as<uint2>(), out.masks.as<uint2>(), total);
  DL_HIP(hipGetLastError());
  voxel_compact(ctx, s, out, e1);
  DL_HIP(hipEventElapsedTime(&out.device_ms, e0, e1));
}

}  // namespace

}  // namespace dl

struct dl_voxel_mesh {
  uint64_t quads = 0;
  int device = 0;
  dl::DeviceBuffer pos, idx;
  float device_ms = 0.0f;
};

using namespace dl;

extern "C" {

int dl_voxel_result_create(dl_context *ctx, const double bounds_min[3], const double bounds_max[3],
                           const double num_blocks[3], double voxel_resolution, const uint32_t *block_idx,
                           const uint32_t *mask_lo, const uint32_t *mask_hi, uint64_t count,
                           dl_voxel_result **out) {
  return guard([&] {
    if (!ctx || !bounds_min || !bounds_max || !num_blocks || !out) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    if (count && (!block_idx || !mask_lo || !mask_hi)) fail(DL_ERR_INVALID, "null block array");
    if (!(voxel_resolution > 0.0) || !std::isfinite(voxel_resolution))
      fail(DL_ERR_INVALID, "voxel_resolution must be positive and finite");
    std::unique_ptr<dl_voxel_result> r(new dl_voxel_result());
    VoxelResult &v = r->r;
    uint64_t total = 1;
    for (int k = 0; k < 3; k++) {
      const double n = num_blocks[k];
      if (!std::isfinite(n) || n < 0.0 || n != std::floor(n) || n > 2147483647.0)
        fail(DL_ERR_INVALID, "num_blocks must be integers in [0, 2^31)");
      v.bmin[k] = bounds_min[k];
      v.bmax[k] = bounds_max[k];
      v.nblocks[k] = n;
      v.nb[k] = (int32_t)n;
      total *= (uint64_t)n;
      if (total >= 0xFFFFFFFFull) fail(DL_ERR_CAPACITY, "voxel grid has more than 2^32-1 blocks");
    }
    v.vres = voxel_resolution;
    for (uint64_t i = 0; i < count; i++)
      if (block_idx[i] >= total) fail(DL_ERR_INVALID, "block index outside the grid");
    Context &c = ctx->c;
    c.make_current();
    v.device = c.device;
    v.total = total;
    hipStream_t s = c.stream;
    if (total) {
      v.masks.reserve(total * 8);
      k_fill_u64<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(v.masks.as<uint64_t>(), total, 0ull);
      DL_HIP(hipGetLastError());
      if (count) {
        DeviceBuffer di, dlo, dhi;
        di.reserve(count * 4);
        dlo.reserve(count * 4);
        dhi.reserve(count * 4);
        DL_HIP(hipMemcpyAsync(di.get(), block_idx, count * 4, hipMemcpyHostToDevice, s));
        DL_HIP(hipMemcpyAsync(dlo.get(), mask_lo, count * 4, hipMemcpyHostToDevice, s));
        DL_HIP(hipMemcpyAsync(dhi.get(), mask_hi, count * 4, hipMemcpyHostToDevice, s));
        k_scatter_blocks<<<(uint32_t)((count + 255) / 256), 256, 0, s>>>(
            di.as<uint32_t>(), dlo.as<uint32_t>(), dhi.as<uint32_t>(), count, v.masks.as<unsigned long long>());
        DL_HIP(hipGetLastError());
        DL_HIP(hipStreamSynchronize(s));
      }
    }
    voxel_compact(c, s, v, nullptr);
    *out = r.release();
  });
}

int dl_voxel_result_device(const dl_voxel_result *res, dl_voxel_device_view *out) {
  return guard([&] {
    if (!res || !out) fail(DL_ERR_INVALID, "null argument");
    const VoxelResult &v = res->r;
    std::memset(out, 0, sizeof(*out));
    out->masks = v.total ? v.masks.as<uint32_t>() : nullptr;
    out->num_blocks = v.total;
    for (int k = 0; k < 3; k++) {
      out->nb[k] = v.nb[k];
      out->bounds_min[k] = v.bmin[k];
    }
    out->voxel_resolution = v.vres;
    out->block_idx = v.count ? v.idx.as<uint32_t>() : nullptr;
    out->mask_lo = v.count ? v.lo.as<uint32_t>() : nullptr;
    out->mask_hi = v.count ? v.hi.as<uint32_t>() : nullptr;
    out->count = v.count;
    out->device = v.device;
  });
}

int dl_voxel_fill(dl_context *ctx, const dl_voxel_result *in, dl_voxel_result **out) {
  return guard([&] {
    if (!ctx || !in || !out) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<dl_voxel_result> r(new dl_voxel_result());
    flood_impl(ctx->c, in->r, true, nullptr, 0, r->r);
    *out = r.release();
  });
}

int dl_voxel_carve(dl_context *ctx, const dl_voxel_result *in, const double *points, uint64_t num_points,
                   dl_voxel_result **out) {
  return guard([&] {
    if (!ctx || !in || !out || (num_points && !points)) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<dl_voxel_result> r(new dl_voxel_result());
    flood_impl(ctx->c, in->r, false, points, num_points, r->r);
    *out = r.release();
  });
}

int dl_voxel_query(dl_context *ctx, const dl_voxel_result *res, const double *points, uint64_t num_points,
                   uint8_t *out) {
  return guard([&] {
    if (!ctx || !res || (num_points && (!points || !out))) fail(DL_ERR_INVALID, "null argument");
    if (!num_points) return;
    const VoxelResult &v = res->r;
    if (v.total == 0) {
      std::memset(out, 0, num_points);
      return;
    }
    Context &c = ctx->c;
    c.make_current();
    hipStream_t s = c.stream;
    DeviceBuffer pts, o;
    pts.reserve(num_points * 24);
    o.reserve(num_points);
    DL_HIP(hipMemcpyAsync(pts.get(), points, num_points * 24, hipMemcpyHostToDevice, s));
    k_query<<<(uint32_t)((num_points + 255) / 256), 256, 0, s>>>(pts.as<double>(), num_points, geom_of(v),
                                                                 v.masks.as<uint2>(), o.as<uint8_t>());
    DL_HIP(hipGetLastError());
    DL_HIP(hipMemcpyAsync(out, o.get(), num_points, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  });
}

int dl_voxel_mesh_build(dl_context *ctx, const dl_voxel_result *res, dl_voxel_mesh **out) {
  return guard([&] {
    if (!ctx || !res || !out) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<dl_voxel_mesh> mesh(new dl_voxel_mesh());
    const VoxelResult &v = res->r;
    Context &c = ctx->c;
    c.make_current();
    mesh->device = c.device;
    hipStream_t s = c.stream;
    hipEvent_t e0, e1;
    DL_HIP(hipEventCreate(&e0));
    DL_HIP(hipEventCreate(&e1));
    struct Events {
      hipEvent_t a, b;
      ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } events{e0, e1};
    DL_HIP(hipEventRecord(e0, s));
    if (v.total == 0 || v.count == 0) {
      DL_HIP(hipEventRecord(e1, s));
      DL_HIP(hipStreamSynchronize(s));
      DL_HIP(hipEventElapsedTime(&mesh->device_ms, e0, e1));
      *out = mesh.release();
      return;
    }
    const GridGeom g = geom_of(v);
    MeshDims m;
    for (int k = 0; k < 3; k++) m.V[k] = 4 * v.nb[k];
    for (int a = 0; a < 3; a++) m.W[a] = (uint32_t)div_up((uint32_t)m.V[kUAxis[a]], 64u);
    m.rowbase[0] = 0;
    m.wordbase[0] = 0;
    for (int d = 0; d < 6; d++) {
      const int a = d >> 1;
      const unsigned long long rows = (unsigned long long)m.V[a] * (unsigned long long)m.V[kVAxis[a]];
      m.rowbase[d + 1] = m.rowbase[d] + rows;
      m.wordbase[d + 1] = m.wordbase[d] + rows * m.W[a];
    }
    const unsigned long long nrows = m.rowbase[6], nwords = m.wordbase[6];
    if (nrows >= 0xFFFFFF00ull || nwords >= 0xFFFFFF00ull) fail(DL_ERR_CAPACITY, "voxel grid too large to mesh");
    DeviceBuffer words, counts, qoff, scan_ws, small;
    words.reserve(nwords * 8);
    counts.reserve(nrows * 4);
    qoff.reserve(nrows * 4);
    scan_ws.reserve(scan_workspace_bytes(nrows));
    small.reserve(256);
    unsigned long long *d_total = reinterpret_cast<unsigned long long *>(small.as<uint8_t>() + 64);
    k_fill_u64<<<1, 256, 0, s>>>(reinterpret_cast<uint64_t *>(d_total), 1, 0ull);
    DL_HIP(hipGetLastError());
    k_face_words<<<(uint32_t)div_up((uint32_t)nwords, 256u), 256, 0, s>>>(v.masks.as<uint2>(), g, m,
                                                                         words.as<unsigned long long>());
    DL_HIP(hipGetLastError());
    const uint32_t row_blocks = (uint32_t)div_up((uint32_t)nrows, 256u);
    k_mesh_rows<false><<<row_blocks, 256, 0, s>>>(words.as<unsigned long long>(), g, m, counts.as<uint32_t>(),
                                                  d_total, nullptr, nullptr, nullptr);
    DL_HIP(hipGetLastError());
    unsigned long long quads = 0;
    DL_HIP(hipMemcpyAsync(&quads, d_total, 8, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    if (quads >= (1ull << 30)) fail(DL_ERR_CAPACITY, "mesh has 2^30 quads or more");
    exclusive_scan_u32(s, counts.as<uint32_t>(), qoff.as<uint32_t>(), nrows, small.as<uint32_t>(), scan_ws.get());
    mesh->quads = quads;
    if (quads) {
      mesh->pos.reserve(quads * 48);
      mesh->idx.reserve(quads * 24);
      k_mesh_rows<true><<<row_blocks, 256, 0, s>>>(words.as<unsigned long long>(), g, m, nullptr, nullptr,
                                                   qoff.as<uint32_t>(), mesh->pos.as<float>(),
                                                   mesh->idx.as<uint32_t>());
      DL_HIP(hipGetLastError());
    }
    DL_HIP(hipEventRecord(e1, s));
    DL_HIP(hipStreamSynchronize(s));
    DL_HIP(hipEventElapsedTime(&mesh->device_ms, e0, e1));
    *out = mesh.release();
  });
}

uint64_t dl_voxel_mesh_quads(const dl_voxel_mesh *m) { return m ? m->quads : 0; }

int dl_voxel_mesh_read(const dl_voxel_mesh *m, float *positions, uint32_t *indices) {
  return guard([&] {
    if (!m || (m->quads && (!positions || !indices))) fail(DL_ERR_INVALID, "null argument");
    if (!m->quads) return;
    DL_HIP(hipSetDevice(m->device));
    DL_HIP(hipMemcpy(positions, m->pos.get(), m->quads * 48, hipMemcpyDeviceToHost));
    DL_HIP(hipMemcpy(indices, m->idx.get(), m->quads * 24, hipMemcpyDeviceToHost));
  });
}

float dl_voxel_mesh_device_ms(const dl_voxel_mesh *m) { return m ? m->device_ms : 0.0f; }

void dl_voxel_mesh_free(dl_voxel_mesh *m) { delete m; }

}  // extern "C"
