// Tile binning without a pair sort (gfx950).
//
// Produces exactly what emit (src/render/shaders/emit-pairs.ts:27-56) followed
// by the stable pair sort (src/render/pipeline.ts:544-578) and the tile
// ranges (src/render/shaders/find-boundaries.ts:17-99) produce: for every tile
// the list of chunk-local splat indices covering it, ascending, plus the
// per-tile offsets. Pair keys are never materialised.
//
// A splat's tiles are a rectangle (a full-row prefix plus one partial row when
// maxCoverage truncates it), so the pair list factors into two stable
// bucketings with few buckets each:
//   level 1  splats -> rows: span (j, x0, x1) for every tile row a splat covers,
//            grouped by row, ascending j inside a row;
//   level 2  spans of one row -> columns: j for every column a span covers,
//            grouped by column, ascending j inside a tile.
// Each level splits its input into fixed blocks (256 splats, 1024 spans). A
// block builds, in LDS, one bitmask per bucket over its elements; popcounts
// give per-(bucket, block) counts, one flat exclusive scan over (bucket,
// block) gives every block's base per bucket, and the rank of an element in
// a bucket is a popcount of the bits below it. Every pair is then written once
// to its final position.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "binning.h"
#include "dl_common.h"
#include "dl_internal.h"

namespace dl {

namespace {

constexpr int kB1 = 256;                 // splats per level-1 block (one per thread)
constexpr int kW1 = kB1 / 32;            // mask words per row
constexpr int kB2Threads = 256;
constexpr int kSpansPerThread = 4;
constexpr int kB2 = kB2Threads * kSpansPerThread;  // spans per level-2 block
constexpr int kW2 = kB2 / 32;            // mask words per column

struct Rect {
  uint32_t x0, x1, y0;
  uint32_t nrows;               // rows y0 .. y0+nrows-1 (0: no pairs)
  uint32_t full_rows;           // rows y0 .. y0+full_rows-1 span [x0, x1]
  uint32_t rem;                 // row y0+full_rows spans [x0, x0+rem-1] when rem > 0
};

// tile rectangle and maxCoverage truncation of one projected splat, with the
// exact arithmetic of project.ts:275-283 / emit-pairs.ts:164-181
__device__ inline Rect splat_rect(const float4 v0, uint32_t tiles_x, uint32_t tiles_y, uint32_t max_cov) {
  Rect R{0, 0, 0, 0, 0, 0};
  const float sx = v0.x, sy = v0.y, radius = v0.z;
  if (!(radius > 0.0f)) return R;
  const uint32_t min_tx = u32sat(wmax(0.0f, __builtin_floorf((sx - radius) / 16.0f)));
  const uint32_t min_ty = u32sat(wmax(0.0f, __builtin_floorf((sy - radius) / 16.0f)));
  const uint32_t max_tx = u32sat(wmin((float)tiles_x - 1.0f, __builtin_floorf((sx + radius) / 16.0f)));
  const uint32_t max_ty = u32sat(wmin((float)tiles_y - 1.0f, __builtin_floorf((sy + radius) / 16.0f)));
  if (max_tx < min_tx || max_ty < min_ty) return R;
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

// ---------------------------- level 1 -------------------------------------
__global__ __launch_bounds__(kB1) void k_rows_count(const float4 *__restrict__ rec, const uint32_t *__restrict__ ids,
                                                    uint32_t cn, BinGrid g, uint32_t nb1, uint32_t *__restrict__ c1) {
  extern __shared__ uint32_t s_mask[];  // [tiles_y][kW1]
  const uint32_t t = threadIdx.x;
  const uint32_t jj = blockIdx.x * kB1 + t;
  for (uint32_t i = t; i < g.tiles_y * kW1; i += kB1) s_mask[i] = 0;
  __syncthreads();
  if (jj < cn) {
    const Rect R = splat_rect(rec[3 * (size_t)ids[jj]], g.tiles_x, g.tiles_y, g.max_cov);
    for (uint32_t k = 0; k < R.nrows; k++) atomicOr(&s_mask[(R.y0 + k) * kW1 + (t >> 5)], 1u << (t & 31));
  }
  __syncthreads();
  for (uint32_t y = t; y < g.tiles_y; y += kB1) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kW1; w++) c += __popc(s_mask[y * kW1 + w]);
    c1[(size_t)y * nb1 + blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(kB1) void k_rows_scatter(const float4 *__restrict__ rec,
                                                      const uint32_t *__restrict__ ids, uint32_t cn, BinGrid g,
                                                      uint32_t nb1, const uint32_t *__restrict__ base1,
                                                      uint2 *__restrict__ spans) {
  extern __shared__ uint32_t s_dyn[];
  uint32_t *s_mask = s_dyn;                                                  // [tiles_y][kW1]
  uint32_t *s_base = s_dyn + g.tiles_y * kW1;                                // [tiles_y]
  uint16_t *s_pre = reinterpret_cast<uint16_t *>(s_base + g.tiles_y);        // [tiles_y][kW1]
  const uint32_t t = threadIdx.x;
  const uint32_t jj = blockIdx.x * kB1 + t;
  for (uint32_t i = t; i < g.tiles_y * kW1; i += kB1) s_mask[i] = 0;
  for (uint32_t y = t; y < g.tiles_y; y += kB1) s_base[y] = base1[(size_t)y * nb1 + blockIdx.x];
  __syncthreads();
  Rect R{0, 0, 0, 0, 0, 0};
  uint32_t id = 0;
  if (jj < cn) {
    id = ids[jj];
    R = splat_rect(rec[3 * (size_t)id], g.tiles_x, g.tiles_y, g.max_cov);
    for (uint32_t k = 0; k < R.nrows; k++) atomicOr(&s_mask[(R.y0 + k) * kW1 + (t >> 5)], 1u << (t & 31));
  }
  __syncthreads();
  for (uint32_t y = t; y < g.tiles_y; y += kB1) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kW1; w++) {
      s_pre[y * kW1 + w] = (uint16_t)run;
      run += __popc(s_mask[y * kW1 + w]);
    }
  }
  __syncthreads();
  if (jj < cn) {
    const uint32_t word = t >> 5, below = (1u << (t & 31)) - 1u;
    for (uint32_t k = 0; k < R.nrows; k++) {
      const uint32_t y = R.y0 + k;
      const uint32_t rank = s_pre[y * kW1 + word] + __popc(s_mask[y * kW1 + word] & below);
      const uint32_t xb = (k < R.full_rows) ? R.x1 : (R.x0 + R.rem - 1u);
      spans[(size_t)s_base[y] + rank] = make_uint2(id, R.x0 | (xb << 16));
    }
  }
}

// --------------------------- row blocks -----------------------------------
// one workgroup: per row, the number of level-2 blocks and their prefix;
// writes row_start[0..tiles_y], block_base[0..tiles_y] and row_of[block]
__global__ __launch_bounds__(1024) void k_row_blocks(const uint32_t *__restrict__ base1, uint32_t nb1,
                                                     uint32_t tiles_y, const uint32_t *__restrict__ d_num_spans,
                                                     uint32_t *__restrict__ row_start,
                                                     uint32_t *__restrict__ block_base,
                                                     uint32_t *__restrict__ row_of, uint32_t max_blocks,
                                                     uint32_t *__restrict__ d_nb2) {
  __shared__ uint32_t s_nb[1024];
  __shared__ uint32_t s_total;
  const uint32_t t = threadIdx.x;
  const uint32_t num_spans = *d_num_spans;
  // tiles_y <= 1024 is checked on the host
  uint32_t start = 0, nb = 0;
  if (t < tiles_y) {
    start = base1[(size_t)t * nb1];
    const uint32_t end = (t + 1 < tiles_y) ? base1[(size_t)(t + 1) * nb1] : num_spans;
    nb = (end - start + kB2 - 1) / kB2;
    row_start[t] = start;
  }
  if (t == 0) row_start[tiles_y] = num_spans;
  s_nb[t] = nb;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 entries
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    const uint32_t v = (t >= d) ? s_nb[t - d] : 0u;
    __syncthreads();
    s_nb[t] += v;
    __syncthreads();
  }
  const uint32_t excl = s_nb[t] - nb;
  if (t < tiles_y) block_base[t] = excl;
  if (t == 1023) {
    s_total = s_nb[1023];
    block_base[tiles_y] = s_total;
    *d_nb2 = s_total;
  }
  __syncthreads();
  if (t < tiles_y)
    for (uint32_t k = 0; k < nb && excl + k < max_blocks; k++) row_of[excl + k] = t;
}

// ---------------------------- level 2 -------------------------------------
struct L2Block {
  uint32_t row, k, nbr, span0, nspan;
  size_t cbase;  // index of (column 0, block k) in the counts array
};

__device__ inline L2Block l2_block(uint32_t g2, const uint32_t *row_of, const uint32_t *row_start,
                                   const uint32_t *block_base, uint32_t tiles_x) {
  L2Block b;
  b.row = row_of[g2];
  const uint32_t bb = block_base[b.row];
  b.k = g2 - bb;
  b.nbr = block_base[b.row + 1] - bb;
  const uint32_t rs = row_start[b.row], re = row_start[b.row + 1];
  b.span0 = rs + b.k * kB2;
  b.nspan = min((uint32_t)kB2, re - b.span0);
  b.cbase = (size_t)tiles_x * bb + b.k;
  return b;
}

__device__ inline void build_col_masks(uint32_t *s_mask, const uint2 *spans, const L2Block &b, uint32_t tiles_x,
                                       uint2 sp[kSpansPerThread]) {
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < tiles_x * kW2; i += kB2Threads) s_mask[i] = 0;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kSpansPerThread; q++) {
    const uint32_t s = t + q * kB2Threads;
    sp[q] = make_uint2(0xFFFFFFFFu, 0u);
    if (s < b.nspan) {
      sp[q] = spans[b.span0 + s];
      const uint32_t xa = sp[q].y & 0xFFFFu, xb = sp[q].y >> 16;
      for (uint32_t x = xa; x <= xb; x++) atomicOr(&s_mask[x * kW2 + (s >> 5)], 1u << (s & 31));
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kB2Threads) void k_cols_count(const uint2 *__restrict__ spans,
                                                           const uint32_t *__restrict__ row_of,
                                                           const uint32_t *__restrict__ row_start,
                                                           const uint32_t *__restrict__ block_base, BinGrid g,
                                                           uint32_t *__restrict__ c2) {
  extern __shared__ uint32_t s_mask[];  // [tiles_x][kW2]
  const L2Block b = l2_block(blockIdx.x, row_of, row_start, block_base, g.tiles_x);
  uint2 sp[kSpansPerThread];
  build_col_masks(s_mask, spans, b, g.tiles_x, sp);
  for (uint32_t x = threadIdx.x; x < g.tiles_x; x += kB2Threads) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kW2; w++) c += __popc(s_mask[x * kW2 + w]);
    c2[b.cbase + (size_t)x * b.nbr] = c;
  }
}

__global__ __launch_bounds__(kB2Threads) void k_cols_scatter(const uint2 *__restrict__ spans,
                                                             const uint32_t *__restrict__ row_of,
                                                             const uint32_t *__restrict__ row_start,
                                                             const uint32_t *__restrict__ block_base, BinGrid g,
                                                             const uint32_t *__restrict__ base2,
                                                             uint32_t *__restrict__ pvals) {
  extern __shared__ uint32_t s_dyn[];
  uint32_t *s_mask = s_dyn;                                                  // [tiles_x][kW2]
  uint32_t *s_base = s_dyn + g.tiles_x * kW2;                                // [tiles_x]
  uint16_t *s_pre = reinterpret_cast<uint16_t *>(s_base + g.tiles_x);        // [tiles_x][kW2]
  const L2Block b = l2_block(blockIdx.x, row_of, row_start, block_base, g.tiles_x);
  for (uint32_t x = threadIdx.x; x < g.tiles_x; x += kB2Threads) s_base[x] = base2[b.cbase + (size_t)x * b.nbr];
  uint2 sp[kSpansPerThread];
  build_col_masks(s_mask, spans, b, g.tiles_x, sp);
  for (uint32_t x = threadIdx.x; x < g.tiles_x; x += kB2Threads) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kW2; w++) {
      s_pre[x * kW2 + w] = (uint16_t)run;
      run += __popc(s_mask[x * kW2 + w]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kSpansPerThread; q++) {
    const uint32_t s = threadIdx.x + q * kB2Threads;
    if (sp[q].x == 0xFFFFFFFFu) continue;
    const uint32_t xa = sp[q].y & 0xFFFFu, xb = sp[q].y >> 16;
    const uint32_t word = s >> 5, below = (1u << (s & 31)) - 1u;
    for (uint32_t x = xa; x <= xb; x++) {
      const uint32_t rank = s_pre[x * kW2 + word] + __popc(s_mask[x * kW2 + word] & below);
      pvals[(size_t)s_base[x] + rank] = sp[q].x;
    }
  }
}

// per-tile pair totals from the (tile, block) counts; rows without spans -> 0
__global__ __launch_bounds__(256) void k_tile_totals(const uint32_t *__restrict__ c2,
                                                     const uint32_t *__restrict__ block_base, BinGrid g,
                                                     uint32_t *__restrict__ totals) {
  const uint32_t tile = blockIdx.x * 256 + threadIdx.x;
  if (tile >= g.tiles_x * g.tiles_y) return;
  const uint32_t row = tile / g.tiles_x, x = tile % g.tiles_x;
  const uint32_t bb = block_base[row], nbr = block_base[row + 1] - bb;
  const size_t base = (size_t)g.tiles_x * bb + (size_t)x * nbr;
  uint32_t s = 0;
  for (uint32_t k = 0; k < nbr; k++) s += c2[base + k];
  totals[tile] = s;
}

}  // namespace

bool span_binning_supported(uint32_t tiles_x, uint32_t tiles_y) {
  return tiles_x >= 1 && tiles_y >= 1 && tiles_x <= kMaxBinTilesX && tiles_y <= kMaxBinTilesY;
}

void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals, uint32_t *d_small) {
  const uint32_t num_tiles = g.tiles_x * g.tiles_y;
  uint32_t *d_num_spans = d_small;
  uint32_t *d_nb2 = d_small + 1;
  uint32_t *d_npairs = d_small + 2;
  // level 1: splats -> row spans (at most one span per pair)
  const uint32_t nb1 = div_up(cn, kB1);
  const size_t n1 = (size_t)g.tiles_y * nb1;
  ws.c1.reserve(n1 * 4);
  ws.base1.reserve(n1 * 4);
  ws.spans.reserve((size_t)std::max<uint32_t>(npairs, 1) * 8);
  ws.scan_ws.reserve(scan_workspace_bytes(n1));
  k_rows_count<<<nb1, kB1, (size_t)g.tiles_y * kW1 * 4, s>>>(rec, ids, cn, g, nb1, ws.c1.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32(s, ws.c1.as<uint32_t>(), ws.base1.as<uint32_t>(), n1, d_num_spans, ws.scan_ws.get());
  const size_t lds1 = (size_t)g.tiles_y * kW1 * 4 + (size_t)g.tiles_y * 4 + (size_t)g.tiles_y * kW1 * 2;
  k_rows_scatter<<<nb1, kB1, lds1, s>>>(rec, ids, cn, g, nb1, ws.base1.as<uint32_t>(), ws.spans.as<uint2>());
  DL_HIP(hipGetLastError());
  const uint32_t max_blocks = div_up(npairs, kB2) + g.tiles_y;
  ws.row_start.reserve(((size_t)g.tiles_y + 1) * 4);
  ws.block_base.reserve(((size_t)g.tiles_y + 1) * 4);
  ws.row_of.reserve((size_t)max_blocks * 4);
  k_row_blocks<<<1, 1024, 0, s>>>(ws.base1.as<uint32_t>(), nb1, g.tiles_y, d_num_spans, ws.row_start.as<uint32_t>(),
                                  ws.block_base.as<uint32_t>(), ws.row_of.as<uint32_t>(), max_blocks, d_nb2);
  DL_HIP(hipGetLastError());
  uint32_t nb2 = 0;
  DL_HIP(hipMemcpyAsync(&nb2, d_nb2, 4, hipMemcpyDeviceToHost, s));
  DL_HIP(hipStreamSynchronize(s));
  if (nb2 > max_blocks) fail(DL_ERR_INTERNAL, "span binning: block table overflow");
  // level 2: row spans -> per-tile lists
  const size_t ncounts = (size_t)g.tiles_x * nb2;
  ws.c2.reserve(std::max<size_t>(ncounts, 1) * 4);
  ws.base2.reserve(std::max<size_t>(ncounts, 1) * 4);
  ws.totals.reserve(((size_t)num_tiles + 1) * 4);
  ws.scan_ws.reserve(scan_workspace_bytes(std::max<size_t>(ncounts, (size_t)num_tiles + 1)));
  if (nb2 > 0) {
    k_cols_count<<<nb2, kB2Threads, (size_t)g.tiles_x * kW2 * 4, s>>>(
        ws.spans.as<uint2>(), ws.row_of.as<uint32_t>(), ws.row_start.as<uint32_t>(), ws.block_base.as<uint32_t>(), g,
        ws.c2.as<uint32_t>());
    DL_HIP(hipGetLastError());
    exclusive_scan_u32(s, ws.c2.as<uint32_t>(), ws.base2.as<uint32_t>(), ncounts, nullptr, ws.scan_ws.get());
  }
  k_tile_totals<<<div_up(num_tiles, 256), 256, 0, s>>>(ws.c2.as<uint32_t>(), ws.block_base.as<uint32_t>(), g,
                                                       ws.totals.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32(s, ws.totals.as<uint32_t>(), d_off, num_tiles, d_npairs, ws.scan_ws.get());
  DL_HIP(hipMemcpyAsync(d_off + num_tiles, d_npairs, 4, hipMemcpyDeviceToDevice, s));
  if (nb2 > 0) {
    const size_t lds2 = (size_t)g.tiles_x * kW2 * 4 + (size_t)g.tiles_x * 4 + (size_t)g.tiles_x * kW2 * 2;
    k_cols_scatter<<<nb2, kB2Threads, lds2, s>>>(ws.spans.as<uint2>(), ws.row_of.as<uint32_t>(),
                                                 ws.row_start.as<uint32_t>(), ws.block_base.as<uint32_t>(), g,
                                                 ws.base2.as<uint32_t>(), pvals);
    DL_HIP(hipGetLastError());
  }
}

}  // namespace dl
