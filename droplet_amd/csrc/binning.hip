// Tile binning without a pair sort (gfx950).
//
// Produces exactly what emit (src/render/shaders/emit-pairs.ts:27-56)
// followed by the stable pair sort (src/render/pipeline.ts:544-578) and the
// tile ranges (src/render/shaders/find-boundaries.ts:17-99) produce: for every
// tile the list of the chunk's splats covering it, in depth order, plus the
// per-tile offsets. Pair keys are never materialised.
//
// A splat's tiles are a rectangle (a full-row prefix plus one partial row
// when maxCoverage truncates it), so the stable pair order factors into
// stable bucketings with few buckets each:
//   level 1  splats -> tile rows: one span (id, x0, x1) per covered row,
//            grouped by row, depth order inside a row (LDS bitmask ranks);
//   level 2a row spans -> 16-column groups of the row;
//   level 2b group spans -> the group's columns: the splat id per tile.
// Levels 2a and 2b are one primitive, a multi-bucket stable scatter (MBS):
// every item covers a contiguous range of at most 16 buckets; a wave ranks its
// 64 items in a bucket with one ballot per bucket, so ranks follow item order;
// per-(segment, bucket, chunk) counts and one flat exclusive scan place every
// chunk's run of each bucket; each item is then written once.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "binning.h"
#include "dl_common.h"
#include "dl_internal.h"

namespace dl {

namespace {

constexpr int kW1 = kB1 / 32;            // mask words per row
constexpr int kChunk = 1024;             // items per MBS chunk: kMbsWaves waves x kMbsSteps steps x 64 lanes
// 4 waves of 4 steps: 2 waves of 8 steps cost ~8 us per 6M-splat 4K frame,
// 8 waves of 2 steps ~60 us
constexpr int kMbsWaves = 4;
constexpr int kMbsSteps = kChunk / (64 * kMbsWaves);
constexpr int kMbsThreads = 64 * kMbsWaves;
constexpr int kK = 16;                   // buckets per MBS segment

// Depth-ordered splat records of one chunk: (id, x0 | x1 << 16, y0 |
// full_rows << 16, rem). One gather of the id-ordered projection records, so
// that both level-1 passes read coalesced 16-byte records.

__device__ inline Rect unpack_rect(const uint4 p) {
  Rect R;
  R.x0 = p.y & 0xFFFFu;
  R.x1 = p.y >> 16;
  R.y0 = p.z & 0xFFFFu;
  R.full_rows = p.z >> 16;
  R.rem = p.w;
  R.nrows = R.full_rows + (R.rem > 0 ? 1u : 0u);
  return R;
}

__global__ __launch_bounds__(256) void k_rects(const float4 *__restrict__ rec, const uint32_t *__restrict__ ids,
                                               uint32_t cn, BinGrid g, uint4 *__restrict__ rects) {
  const uint32_t jj = blockIdx.x * 256 + threadIdx.x;
  if (jj >= cn) return;
  const uint32_t id = ids[jj];
  rects[jj] = pack_rect(id, splat_rect(rec[3 * (size_t)id], g.tg, g.max_cov));
}

// ------------------------ live-splat compaction ----------------------------
// In a pass after the first, a splat enters level 1 only if some row of its
// rect holds a live tile in the row's column span: exactly the splats that
// would give level 1 at least one span. Removing the others from the
// depth-ordered input (a stable compaction) leaves every list unchanged.

// Summed-area table of the open tiles from the row prefixes: sat[y * (cols
// + 1) + x] = open tiles in rows < y and columns < x, for y <= rows, x <=
// cols. One thread per column accumulates down the rows. Then csat, the
// table at the corners of 4x4-tile blocks (LiveMap::csat).
__global__ __launch_bounds__(256) void k_open_sat(const uint16_t *__restrict__ pre, uint32_t rows, uint32_t cols,
                                                  uint32_t *__restrict__ sat, uint32_t *__restrict__ csat) {
  const uint32_t stride = cols + 1;
  for (uint32_t x = threadIdx.x; x <= cols; x += 256) {
    uint32_t acc = 0;
    sat[x] = 0;
#pragma unroll 8
    for (uint32_t y = 0; y < rows; y++) {
      acc += pre[(size_t)y * stride + x];
      sat[(size_t)(y + 1) * stride + x] = acc;
    }
  }
  __syncthreads();
  const uint32_t cstride = (cols + 3) / 4 + 1, crows = (rows + 3) / 4 + 1;
  for (uint32_t k = threadIdx.x; k < crows * cstride; k += 256) {
    const uint32_t cy = k / cstride, cx = k % cstride;
    csat[k] = sat[(size_t)min(4 * cy, rows) * stride + min(4 * cx, cols)];
  }
}

// The same table for a tall stacked image (a batch group's views): one
// workgroup per column x <= cols scans that column down the rows (a single
// workgroup walking thousands of rows per column would take a millisecond),
// then k_open_csat samples the 4x4-block corners.
__global__ __launch_bounds__(256) void k_open_sat_cols(const uint16_t *__restrict__ pre, uint32_t rows, uint32_t cols,
                                                       uint32_t *__restrict__ sat) {
  __shared__ uint32_t s_w[4];
  const uint32_t x = blockIdx.x, stride = cols + 1;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) sat[x] = 0;
  uint32_t carry = 0;
  for (uint32_t y0 = 0; y0 < rows; y0 += 256) {
    const uint32_t y = y0 + threadIdx.x;
    const uint32_t v = y < rows ? (uint32_t)pre[(size_t)y * stride + x] : 0u;
    uint32_t inc = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, d, 64);
      if (lane >= d) inc += t;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t before = carry;
    for (uint32_t w = 0; w < wave; w++) before += s_w[w];
    if (y < rows) sat[(size_t)(y + 1) * stride + x] = before + inc;
    carry += s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_open_csat(const uint32_t *__restrict__ sat, uint32_t rows, uint32_t cols,
                                                   uint32_t *__restrict__ csat) {
  const uint32_t stride = cols + 1, cstride = (cols + 3) / 4 + 1, crows = (rows + 3) / 4 + 1;
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= crows * cstride) return;
  const uint32_t cy = k / cstride, cx = k % cstride;
  csat[k] = sat[(size_t)min(4 * cy, rows) * stride + min(4 * cx, cols)];
}

// The compaction kernels give each thread kLiveItems items (one per 256-item
// sub-block, loads issued together): the per-item work is a few loads, so
// one item per thread leaves them latency-bound.
constexpr int kLiveItems = 8;

// bit i of bits: splat i (id order, packed rects of the projection) is live
// (rect_live)
__global__ __launch_bounds__(256) void k_live_bits(const uint2 *__restrict__ rect8, uint32_t n, LiveMap live,
                                                   unsigned long long *__restrict__ bits) {
  const size_t base = (size_t)blockIdx.x * 256 * kLiveItems + threadIdx.x;
  uint2 r[kLiveItems];
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    const size_t i = base + (size_t)k * 256;
    r[k] = i < n ? rect8[i] : make_uint2(0u, 0u);
  }
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    const size_t i = base + (size_t)k * 256;
    const bool lv = rect_live(unpack_rect8(r[k]), live);
    const unsigned long long b = __ballot(lv);
    const size_t first = i - (size_t)__lane_id();
    if (__lane_id() == 0 && first < n) bits[first >> 6] = b;
  }
}

__device__ inline bool live_bit(const unsigned long long *__restrict__ bits, uint32_t id) {
  return (bits[id >> 6] >> (id & 63u)) & 1ull;
}

// Is splat id live: its bit, or without bits (a pass of few splats) the
// test of its own rect
__device__ inline bool splat_live(const unsigned long long *__restrict__ bits, const uint2 *__restrict__ rect8,
                                  const LiveMap &lm, uint32_t id) {
  return bits ? live_bit(bits, id) : rect_live(unpack_rect8(rect8[id]), lm);
}

// live splats per workgroup of kLiveItems 256-item sub-blocks of the pass
// (from the live bits, or with bits null from each splat's rect)
__global__ __launch_bounds__(256) void k_live_count(const uint32_t *__restrict__ ids, uint32_t cn,
                                                    const unsigned long long *__restrict__ bits,
                                                    const uint2 *__restrict__ rect8, LiveMap lm,
                                                    uint32_t *__restrict__ blk_tot) {
  __shared__ uint32_t s_w[kLiveItems][4];
  const uint32_t sub0 = blockIdx.x * kLiveItems;
  uint32_t id[kLiveItems];
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    const uint32_t jj = (sub0 + k) * 256 + threadIdx.x;
    id[k] = jj < cn ? ids[jj] : 0xFFFFFFFFu;
  }
  bool lv[kLiveItems];
  if (bits) {
    unsigned long long w[kLiveItems];
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) w[k] = id[k] != 0xFFFFFFFFu ? bits[id[k] >> 6] : 0ull;
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) lv[k] = id[k] != 0xFFFFFFFFu && ((w[k] >> (id[k] & 63u)) & 1ull);
  } else {
    uint2 r[kLiveItems];
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) r[k] = id[k] != 0xFFFFFFFFu ? rect8[id[k]] : make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) lv[k] = id[k] != 0xFFFFFFFFu && rect_live(unpack_rect8(r[k]), lm);
  }
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    const unsigned long long b = __ballot(lv[k]);
    if (__lane_id() == 0) s_w[k][threadIdx.x >> 6] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) t += s_w[k][0] + s_w[k][1] + s_w[k][2] + s_w[k][3];
    blk_tot[blockIdx.x] = t;
  }
}

// the pass's live splats as packed depth-ordered rects, compacted: each
// workgroup sums the totals of the workgroups before it (k_live_count) for
// its start; the last one writes live_small[0] = live items and
// live_small[1] = their level-1 blocks
__global__ __launch_bounds__(256) void k_rects_live(const uint2 *__restrict__ rect8, const uint32_t *__restrict__ ids,
                                                    uint32_t cn, const unsigned long long *__restrict__ bits,
                                                    LiveMap lm, const uint32_t *__restrict__ blk_tot,
                                                    uint32_t *__restrict__ live_small, uint4 *__restrict__ rects) {
  __shared__ uint32_t s_w[kLiveItems][4];
  __shared__ uint32_t s_pre[4];
  const uint32_t sub0 = blockIdx.x * kLiveItems;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  uint32_t id[kLiveItems];
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    const uint32_t jj = (sub0 + k) * 256 + threadIdx.x;
    id[k] = jj < cn ? ids[jj] : 0xFFFFFFFFu;
  }
  unsigned long long bal[kLiveItems];
  uint2 r[kLiveItems];
  if (bits) {
    unsigned long long w[kLiveItems];
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) w[k] = id[k] != 0xFFFFFFFFu ? bits[id[k] >> 6] : 0ull;
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) {
      const bool lv = id[k] != 0xFFFFFFFFu && ((w[k] >> (id[k] & 63u)) & 1ull);
      bal[k] = __ballot(lv);
      r[k] = lv ? rect8[id[k]] : make_uint2(0u, 0u);
    }
  } else {
    // a pass of few splats: each splat's own rect, tested here
#pragma unroll
    for (int k = 0; k < kLiveItems; k++) r[k] = id[k] != 0xFFFFFFFFu ? rect8[id[k]] : make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < kLiveItems; k++)
      bal[k] = __ballot(id[k] != 0xFFFFFFFFu && rect_live(unpack_rect8(r[k]), lm));
  }
#pragma unroll
  for (int k = 0; k < kLiveItems; k++)
    if (lane == 0) s_w[k][wave] = (uint32_t)__popcll(bal[k]);
  // live items of the workgroups before this one
  uint32_t pre = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256) pre += blk_tot[b];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) pre += (uint32_t)__shfl_xor((int)pre, d, 64);
  if (lane == 0) s_pre[wave] = pre;
  __syncthreads();
  uint32_t start = s_pre[0] + s_pre[1] + s_pre[2] + s_pre[3];
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
  for (int k = 0; k < kLiveItems; k++) {
    if ((bal[k] >> lane) & 1ull) {
      uint32_t pos = start + (uint32_t)__popcll(bal[k] & below);
      for (int q = 0; q < wave; q++) pos += s_w[k][q];
      rects[pos] = pack_rect(id[k], unpack_rect8(r[k]));
    }
    start += s_w[k][0] + s_w[k][1] + s_w[k][2] + s_w[k][3];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    live_small[0] = start;
    live_small[1] = (start + kB1 - 1) / kB1;
  }
}

// The level-1 view that owns level-1 block b: the single view `one`, or a
// binary search of the batch table by first block.
__device__ inline L1View l1_view(const L1View &one, const L1View *__restrict__ tab, uint32_t ntab, uint32_t b) {
  if (!tab) return one;
  uint32_t lo = 0, hi = ntab - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (tab[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return tab[lo];
}

__global__ __launch_bounds__(kB1) void k_batch_rects(const float4 *__restrict__ rec, const uint32_t *__restrict__ ids,
                                                     const L1View *__restrict__ tab, uint32_t ntab,
                                                     uint4 *__restrict__ rects) {
  const L1View v = l1_view(L1View{}, tab, ntab, blockIdx.x);
  const uint32_t jj = (blockIdx.x - v.blk0) * kB1 + threadIdx.x;
  if (jj >= v.n) return;
  const uint32_t id = ids[v.item0 + jj];
  rects[v.item0 + jj] = pack_rect(id, splat_rect(rec[3 * (size_t)id], grid16(v.tiles_x, v.rows), v.max_cov));
}

// ---------------------------- level 1 -------------------------------------
// Level-1 block b of view v covers items [item0 + 256 k, ...) with k = b -
// blk0, and rows [row0, row0 + rows); its count of row r sits at c1_base +
// (r - row0) * nb1 + k, so the flat scan orders (view, row, block).
// With d_items (single view) the view's item count is read on the device:
// the grid covers the capacity and blocks past the count exit at once.
__device__ inline L1View l1_view_dev(const L1View &one, const L1View *__restrict__ tab, uint32_t ntab, uint32_t b,
                                     const uint32_t *__restrict__ d_items) {
  L1View v = l1_view(one, tab, ntab, b);
  if (d_items) {
    v.n = min(*d_items, v.n);
    v.n_live = v.n;
  }
  return v;
}

__global__ __launch_bounds__(kB1) void k_rows_count(const uint4 *__restrict__ rects, L1View one,
                                                    const L1View *__restrict__ tab, uint32_t ntab, LiveMap live,
                                                    const uint32_t *__restrict__ d_items,
                                                    uint32_t *__restrict__ c1) {
  extern __shared__ uint32_t s_mask[];  // [rows][kW1]
  const L1View v = l1_view_dev(one, tab, ntab, blockIdx.x, d_items);
  const uint32_t t = threadIdx.x;
  const uint32_t k_blk = blockIdx.x - v.blk0;
  const uint32_t nb1 = (v.n + kB1 - 1) / kB1;
  if (k_blk >= nb1) return;
  if (k_blk * kB1 >= v.n_live) {
    // past the items read (a compacted pass): zero counts
    for (uint32_t y = t; y < v.rows; y += kB1) c1[(size_t)v.c1_base + (size_t)y * nb1 + k_blk] = 0;
    return;
  }
  const uint32_t jj = k_blk * kB1 + t;
  for (uint32_t i = t; i < v.rows * kW1; i += kB1) s_mask[i] = 0;
  __syncthreads();
  if (jj < v.n_live) {
    const Rect R = unpack_rect(rects[v.item0 + jj]);
    for (uint32_t k = 0; k < R.nrows; k++) {
      const uint32_t xb = (k < R.full_rows) ? R.x1 : (R.x0 + R.rem - 1u);
      if (live_range(live, v.row0 + R.y0 + k, R.x0, xb))
        atomicOr(&s_mask[(R.y0 + k) * kW1 + (t >> 5)], 1u << (t & 31));
    }
  }
  __syncthreads();
  for (uint32_t y = t; y < v.rows; y += kB1) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kW1; w++) c += __popc(s_mask[y * kW1 + w]);
    c1[(size_t)v.c1_base + (size_t)y * nb1 + k_blk] = c;
  }
}

__global__ __launch_bounds__(kB1) void k_rows_scatter(const uint4 *__restrict__ rects, L1View one,
                                                      const L1View *__restrict__ tab, uint32_t ntab, LiveMap live,
                                                      const uint32_t *__restrict__ d_items,
                                                      const uint32_t *__restrict__ base1, uint32_t cap,
                                                      uint2 *__restrict__ spans) {
  extern __shared__ uint32_t s_dyn[];
  const L1View v = l1_view_dev(one, tab, ntab, blockIdx.x, d_items);
  uint32_t *s_mask = s_dyn;                                                  // [rows][kW1]
  uint32_t *s_base = s_dyn + v.rows * kW1;                                   // [rows]
  uint16_t *s_pre = reinterpret_cast<uint16_t *>(s_base + v.rows);           // [rows][kW1]
  const uint32_t t = threadIdx.x;
  const uint32_t k_blk = blockIdx.x - v.blk0;
  const uint32_t nb1 = (v.n + kB1 - 1) / kB1;
  if (k_blk >= nb1 || k_blk * kB1 >= v.n_live) return;
  const uint32_t jj = k_blk * kB1 + t;
  for (uint32_t i = t; i < v.rows * kW1; i += kB1) s_mask[i] = 0;
  for (uint32_t y = t; y < v.rows; y += kB1) s_base[y] = base1[(size_t)v.c1_base + (size_t)y * nb1 + k_blk];
  __syncthreads();
  Rect R{0, 0, 0, 0, 0, 0};
  uint32_t id = 0;
  uint32_t rows_live = 0;  // bit k: row R.y0 + k has a live tile in the span (nrows <= 32 rows tracked)
  if (jj < v.n_live) {
    const uint4 pr = rects[v.item0 + jj];
    id = pr.x;
    R = unpack_rect(pr);
    for (uint32_t k = 0; k < R.nrows; k++) {
      const uint32_t xb = (k < R.full_rows) ? R.x1 : (R.x0 + R.rem - 1u);
      if (live_range(live, v.row0 + R.y0 + k, R.x0, xb)) {
        atomicOr(&s_mask[(R.y0 + k) * kW1 + (t >> 5)], 1u << (t & 31));
        if (k < 32) rows_live |= 1u << k;
      }
    }
  }
  __syncthreads();
  for (uint32_t y = t; y < v.rows; y += kB1) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kW1; w++) {
      s_pre[y * kW1 + w] = (uint16_t)run;
      run += __popc(s_mask[y * kW1 + w]);
    }
  }
  __syncthreads();
  if (jj < v.n_live) {
    const uint32_t word = t >> 5, below = (1u << (t & 31)) - 1u;
    for (uint32_t k = 0; k < R.nrows; k++) {
      const uint32_t xb = (k < R.full_rows) ? R.x1 : (R.x0 + R.rem - 1u);
      if (k < 32 ? !((rows_live >> k) & 1u) : !live_range(live, v.row0 + R.y0 + k, R.x0, xb)) continue;
      const uint32_t y = R.y0 + k;
      const uint32_t rank = s_pre[y * kW1 + word] + __popc(s_mask[y * kW1 + word] & below);
      const uint32_t pos = s_base[y] + rank;
      if (pos < cap) spans[pos] = make_uint2(id, R.x0 | (xb << 16));
    }
  }
}

// row_start[r] = spans before virtual row r (the (row r, block 0) entry of
// the level-1 scan, or the running total for a view without blocks),
// row_start[rows] = all spans; clamped to the span capacity cap
__device__ inline void row_start_one(uint32_t r, const uint32_t *__restrict__ base1, size_t n1, const L1View &one,
                                     const L1View *__restrict__ tab, uint32_t ntab, uint32_t rows,
                                     const uint32_t *__restrict__ d_items, const uint32_t *__restrict__ d_num_spans,
                                     uint32_t cap, uint32_t *__restrict__ row_start) {
  const uint32_t total = min(*d_num_spans, cap);
  if (r == rows) row_start[r] = total;
  if (r >= rows) return;
  L1View v = one;
  if (tab) {
    uint32_t lo = 0, hi = ntab - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (tab[mid].row0 <= r) lo = mid; else hi = mid - 1;
    }
    v = tab[lo];
  }
  if (d_items) v.n = min(*d_items, v.n);
  const uint32_t nb1 = (v.n + kB1 - 1) / kB1;
  // with a device count the level-1 counts fill rows * nb1 entries
  const size_t n1e = d_items ? (size_t)v.rows * nb1 : n1;
  const size_t idx = (size_t)v.c1_base + (size_t)(r - v.row0) * nb1;
  row_start[r] = idx < n1e ? min(base1[idx], cap) : total;
}

__global__ __launch_bounds__(256) void k_row_start(const uint32_t *__restrict__ base1, size_t n1, L1View one,
                                                   const L1View *__restrict__ tab, uint32_t ntab, uint32_t rows,
                                                   const uint32_t *__restrict__ d_items,
                                                   const uint32_t *__restrict__ d_num_spans, uint32_t cap,
                                                   uint32_t *__restrict__ row_start) {
  row_start_one(blockIdx.x * 256 + threadIdx.x, base1, n1, one, tab, ntab, rows, d_items, d_num_spans, cap,
                row_start);
}

// ----------------------- MBS chunk tables ---------------------------------
__global__ __launch_bounds__(256) void k_seg_chunks(const uint32_t *__restrict__ seg_start, uint32_t nseg,
                                                    uint32_t *__restrict__ nchunks) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s < nseg) nchunks[s] = (seg_start[s + 1] - seg_start[s] + kChunk - 1) / kChunk;
}

__global__ __launch_bounds__(256) void k_seg_of(const uint32_t *__restrict__ chunk_base, uint32_t nseg, uint32_t cap,
                                                uint32_t *__restrict__ seg_of) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nseg) return;
  const uint32_t e = min(chunk_base[s + 1], cap);
  for (uint32_t c = chunk_base[s]; c < e; c++) seg_of[c] = s;
}

// ---------------------------- MBS -----------------------------------------
struct MbsGeom {
  const uint32_t *seg_of;      // chunk -> segment
  const uint32_t *chunk_base;  // segment -> first chunk (nseg + 1 entries; [nseg] = chunks)
  const uint32_t *seg_start;   // segment -> first item (nseg + 1 entries)
  uint32_t groups;             // 16-column groups per row (level 2b: segment = row * groups + group)
  uint32_t nseg;
  uint32_t chunk_cap;          // capacity of the chunk table
  uint32_t out_cap;            // capacity of the output array
  LiveMap live;                // tiles still open (no map: all)
};

// chunks of this level (read on the device), clamped to the table capacity
__device__ inline uint32_t mbs_chunks(const MbsGeom &G) { return min(G.chunk_base[G.nseg], G.chunk_cap); }

// Buckets of an item as a 16-bit mask: the groups (level 0, segment = row)
// or the columns of the segment's group (level 1, segment = row * groups +
// group) that the item's column range [xa, xb] covers and that hold a live
// tile.
template <int LEVEL>
__device__ inline uint32_t item_buckets(uint2 it, uint32_t seg, const MbsGeom &G) {
  const uint32_t xa = it.y & 0xFFFFu, xb = it.y >> 16;
  uint32_t lo, hi, row, g0;
  if (LEVEL == 0) {
    lo = xa >> 4;
    hi = xb >> 4;
    row = seg;
    g0 = 0;
  } else {
    g0 = (seg % G.groups) * 16u;
    lo = max(xa, g0) - g0;
    hi = min(xb, g0 + 15u) - g0;
    row = seg / G.groups;
  }
  if (lo > hi) return 0;
  const uint32_t span = ((hi - lo + 1u) >= 32u ? 0xFFFFFFFFu : ((1u << (hi - lo + 1u)) - 1u)) << lo;
  if (!G.live.pre) return span & 0xFFFFu;
  uint32_t m = 0;
  for (uint32_t b = lo; b <= hi; b++) {
    const uint32_t c0 = LEVEL == 0 ? max(xa, 16u * b) : g0 + b;
    const uint32_t c1 = LEVEL == 0 ? min(xb, 16u * b + 15u) : g0 + b;
    if (live_range(G.live, row, c0, c1)) m |= 1u << b;
  }
  return m;
}

struct ChunkSpan {
  uint32_t seg, cb, nch, k, i0, iend;
};

__device__ inline ChunkSpan chunk_span(const MbsGeom &G, uint32_t c) {
  ChunkSpan s;
  s.seg = G.seg_of[c];
  s.cb = G.chunk_base[s.seg];
  s.nch = G.chunk_base[s.seg + 1] - s.cb;
  s.k = c - s.cb;
  s.i0 = G.seg_start[s.seg] + s.k * kChunk;
  s.iend = min(G.seg_start[s.seg + 1], s.i0 + kChunk);
  return s;
}

// counts[(16 * cb + b) * ... ] laid out per segment as [bucket][chunk]:
// index 16*cb + b*nch + k, so a flat exclusive scan orders (segment, bucket, chunk)
// Both MBS kernels run a fixed grid that strides over the chunks, whose
// number is only known on the device.
template <int LEVEL>
__global__ __launch_bounds__(kMbsThreads) void k_mbs_count(const uint2 *__restrict__ items, MbsGeom G,
                                                           uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_cnt[kMbsWaves][kK];
  const uint32_t nch = mbs_chunks(G);
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
  const ChunkSpan cs = chunk_span(G, c);
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t cnt[kK];
#pragma unroll
  for (int b = 0; b < kK; b++) cnt[b] = 0;
#pragma unroll
  for (int step = 0; step < kMbsSteps; step++) {
    const uint32_t i = cs.i0 + (uint32_t)w * (64u * kMbsSteps) + (uint32_t)step * 64u + (uint32_t)lane;
    const uint32_t bm = i < cs.iend ? item_buckets<LEVEL>(items[i], cs.seg, G) : 0u;
#pragma unroll
    for (int b = 0; b < kK; b++) cnt[b] += (uint32_t)__popcll(__ballot((bm >> b) & 1u));
  }
  if (lane == 0) {
#pragma unroll
    for (int b = 0; b < kK; b++) s_cnt[w][b] = cnt[b];
  }
  __syncthreads();
  if (threadIdx.x < kK) {
    const uint32_t b = threadIdx.x;
    uint32_t t = 0;
#pragma unroll
    for (int v = 0; v < kMbsWaves; v++) t += s_cnt[v][b];
    counts[(size_t)kK * cs.cb + (size_t)b * cs.nch + cs.k] = t;
  }
  __syncthreads();
  }
}

template <int LEVEL>
__global__ __launch_bounds__(kMbsThreads) void k_mbs_write(const uint2 *__restrict__ items, MbsGeom G,
                                                           const uint32_t *__restrict__ base,
                                                           uint2 *__restrict__ out_items,
                                                           uint32_t *__restrict__ out_ids) {
  __shared__ uint32_t s_cnt[kMbsWaves][kK];
  const uint32_t nch = mbs_chunks(G);
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
  const ChunkSpan cs = chunk_span(G, c);
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint2 it[kMbsSteps];
  uint32_t bm[kMbsSteps];
  uint32_t cnt[kK];
#pragma unroll
  for (int b = 0; b < kK; b++) cnt[b] = 0;
#pragma unroll
  for (int step = 0; step < kMbsSteps; step++) {
    const uint32_t i = cs.i0 + (uint32_t)w * (64u * kMbsSteps) + (uint32_t)step * 64u + (uint32_t)lane;
    bm[step] = 0;
    it[step] = make_uint2(0u, 0u);
    if (i < cs.iend) {
      it[step] = items[i];
      bm[step] = item_buckets<LEVEL>(it[step], cs.seg, G);
    }
#pragma unroll
    for (int b = 0; b < kK; b++) cnt[b] += (uint32_t)__popcll(__ballot((bm[step] >> b) & 1u));
  }
  if (lane == 0) {
#pragma unroll
    for (int b = 0; b < kK; b++) s_cnt[w][b] = cnt[b];
  }
  __syncthreads();
  uint32_t run[kK];
#pragma unroll
  for (int b = 0; b < kK; b++) {
    uint32_t r = base[(size_t)kK * cs.cb + (size_t)b * cs.nch + cs.k];
    for (int v = 0; v < w; v++) r += s_cnt[v][b];
    run[b] = r;
  }
#pragma unroll
  for (int step = 0; step < kMbsSteps; step++) {
#pragma unroll
    for (int b = 0; b < kK; b++) {
      const bool in = (bm[step] >> b) & 1u;
      const uint64_t m = __ballot(in);
      if (in) {
        const uint32_t pos = run[b] + (uint32_t)__popcll(m & lt);
        if (pos < G.out_cap) {
          if (LEVEL == 0)
            out_items[pos] = it[step];
          else
            out_ids[pos] = it[step].x;
        }
      }
      run[b] += (uint32_t)__popcll(m);
    }
  }
  __syncthreads();
  }
}

// Start of every (segment, bucket) list from the scanned MBS counts: the
// (segment, bucket, chunk 0) entry; a segment without chunks takes the entry
// where it would start, or the total at the end. Level 2a: o = row*groups +
// group. Level 2b: o = tile = row*tiles_x + x -> (row*groups + x/16, x%16).
__device__ inline void bucket_offset_one(uint32_t o, const uint32_t *__restrict__ base,
                                         const uint32_t *__restrict__ chunk_base, uint32_t nseg, uint32_t chunk_cap,
                                         uint32_t n_out, uint32_t groups, uint32_t tiles_x, int tile_mode,
                                         const uint32_t *__restrict__ d_total, uint32_t out_cap,
                                         uint32_t *__restrict__ out, unsigned long long *__restrict__ acc) {
  if (o > n_out) return;
  const uint32_t total = min(*d_total, out_cap);
  if (o == n_out) {
    out[o] = total;
    if (acc) *acc += *d_total;
    return;
  }
  const size_t ncounts = (size_t)kK * min(chunk_base[nseg], chunk_cap);
  uint32_t seg, b;
  if (tile_mode) {
    const uint32_t r = o / tiles_x, x = o % tiles_x;
    seg = r * groups + x / 16u;
    b = x % 16u;
  } else {
    seg = o / groups;
    b = o % groups;
  }
  const uint32_t cb = chunk_base[seg], nch = chunk_base[seg + 1] - cb;
  const size_t pos = (size_t)kK * cb + (size_t)b * nch;
  out[o] = pos < ncounts ? min(base[pos], total) : total;
}

__global__ __launch_bounds__(256) void k_bucket_offsets(const uint32_t *__restrict__ base,
                                                        const uint32_t *__restrict__ chunk_base, uint32_t nseg,
                                                        uint32_t chunk_cap, uint32_t n_out, uint32_t groups,
                                                        uint32_t tiles_x, int tile_mode,
                                                        const uint32_t *__restrict__ d_total, uint32_t out_cap,
                                                        uint32_t *__restrict__ out,
                                                        unsigned long long *__restrict__ acc) {
  bucket_offset_one(blockIdx.x * 256 + threadIdx.x, base, chunk_base, nseg, chunk_cap, n_out, groups, tiles_x,
                    tile_mode, d_total, out_cap, out, acc);
}

// The whole chunk table of a level with at most kSmallSegs segments in one
// 1024-thread workgroup: each thread owns 4 consecutive segments.
constexpr uint32_t kSmallSegs = 4096;

// The chunk bases also go to s_cb (LDS, nseg + 1 entries), and then every
// thread fills seg_of for chunks t, t + 1024, ... by a binary search of
// s_cb (a thread filling its own segments' chunks left a few threads
// writing hundreds of entries each: 11.9 us for the 68 rows of a 4K pass).
__device__ inline void chunk_table_small(const uint32_t *__restrict__ seg_start, uint32_t nseg, uint32_t cap,
                                         uint32_t *__restrict__ chunk_base, uint32_t *__restrict__ seg_of,
                                         uint32_t *s_wave, uint32_t *s_cb) {
  const uint32_t t = threadIdx.x, s0 = 4 * t;
  uint32_t nch[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t sg = s0 + k;
    nch[k] = sg < nseg ? (seg_start[sg + 1] - seg_start[sg] + kChunk - 1) / kChunk : 0u;
    sum += nch[k];
  }
  // exclusive scan of the per-thread sums over the 16 waves
  const int lane = __lane_id(), wave = t >> 6;
  uint32_t inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t run = inc - sum, total = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wave) run += s_wave[w];
    total += s_wave[w];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t sg = s0 + k;
    if (sg < nseg) {
      chunk_base[sg] = run;
      s_cb[sg] = run;
    }
    run += nch[k];
  }
  if (t == 0) chunk_base[nseg] = total;
  __syncthreads();
  // chunk c belongs to the last segment whose base is <= c (a segment
  // without chunks shares the next one's base and is passed over)
  const uint32_t e = min(total, cap);
  for (uint32_t c = t; c < e; c += 1024) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (s_cb[mid] <= c) lo = mid; else hi = mid - 1;
    }
    seg_of[c] = lo;
  }
}

__global__ __launch_bounds__(1024) void k_chunk_table_small(const uint32_t *__restrict__ seg_start, uint32_t nseg,
                                                            uint32_t cap, uint32_t *__restrict__ chunk_base,
                                                            uint32_t *__restrict__ seg_of) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_cb[kSmallSegs];
  chunk_table_small(seg_start, nseg, cap, chunk_base, seg_of, s_wave, s_cb);
}

// Level 1's row starts, then level 2a's chunk table over them, in one
// workgroup (rows <= kSmallSegs): the barrier makes the row starts visible.
__global__ __launch_bounds__(1024) void k_row_start_chunks(const uint32_t *__restrict__ base1, size_t n1, L1View one,
                                                          const L1View *__restrict__ tab, uint32_t ntab,
                                                          uint32_t rows, const uint32_t *__restrict__ d_items,
                                                          const uint32_t *__restrict__ d_num_spans, uint32_t cap,
                                                          uint32_t *__restrict__ row_start, uint32_t max_chunks,
                                                          uint32_t *__restrict__ chunk_base,
                                                          uint32_t *__restrict__ seg_of) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_cb[kSmallSegs];
  for (uint32_t r = threadIdx.x; r <= rows; r += 1024)
    row_start_one(r, base1, n1, one, tab, ntab, rows, d_items, d_num_spans, cap, row_start);
  __syncthreads();
  chunk_table_small(row_start, rows, max_chunks, chunk_base, seg_of, s_wave, s_cb);
}

// Level 2a's group offsets (the segment starts of level 2b), then level 2b's
// chunk table over them, in one workgroup (segments <= kSmallSegs).
__global__ __launch_bounds__(1024) void k_offsets_chunks(const uint32_t *__restrict__ base,
                                                        const uint32_t *__restrict__ chunk_base_a, uint32_t rows,
                                                        uint32_t max_ch_a, uint32_t nseg_b, uint32_t groups,
                                                        uint32_t tiles_x, const uint32_t *__restrict__ d_total,
                                                        uint32_t cap, uint32_t *__restrict__ seg_start_b,
                                                        uint32_t max_ch_b, uint32_t *__restrict__ chunk_base_b,
                                                        uint32_t *__restrict__ seg_of_b) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_cb[kSmallSegs];
  for (uint32_t o = threadIdx.x; o <= nseg_b; o += 1024)
    bucket_offset_one(o, base, chunk_base_a, rows, max_ch_a, nseg_b, groups, tiles_x, 0, d_total, cap, seg_start_b,
                      nullptr);
  __syncthreads();
  chunk_table_small(seg_start_b, nseg_b, max_ch_b, chunk_base_b, seg_of_b, s_wave, s_cb);
}

// Chunk table of one MBS level, built on the device: chunks per segment,
// their scan (chunk_base, with the total in chunk_base[nseg]) and chunk ->
// segment. max_chunks bounds the total: every segment of len items has
// ceil(len / kChunk) <= len / kChunk + 1 chunks.
void build_chunk_table(hipStream_t s, const uint32_t *seg_start, uint32_t nseg, DeviceBuffer &tmp,
                       DeviceBuffer &chunk_base, DeviceBuffer &seg_of, DeviceBuffer &scan_ws, uint32_t max_chunks) {
  tmp.reserve((size_t)nseg * 4 + 4);
  chunk_base.reserve(((size_t)nseg + 1) * 4);
  seg_of.reserve((size_t)std::max<uint32_t>(max_chunks, 1) * 4);
  if (nseg <= kSmallSegs) {
    k_chunk_table_small<<<1, 1024, 0, s>>>(seg_start, nseg, max_chunks, chunk_base.as<uint32_t>(),
                                           seg_of.as<uint32_t>());
    DL_HIP(hipGetLastError());
    return;
  }
  scan_ws.reserve(scan_workspace_bytes(nseg));
  k_seg_chunks<<<div_up(nseg, 256), 256, 0, s>>>(seg_start, nseg, tmp.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32(s, tmp.as<uint32_t>(), chunk_base.as<uint32_t>(), nseg, chunk_base.as<uint32_t>() + nseg,
                     scan_ws.get());
  k_seg_of<<<div_up(nseg, 256), 256, 0, s>>>(chunk_base.as<uint32_t>(), nseg, max_chunks, seg_of.as<uint32_t>());
  DL_HIP(hipGetLastError());
}

// Fixed grid of the chunk-striding MBS kernels. The kernels wait on their
// loads more than they compute, so the grid is large enough to give most
// chunks a workgroup of their own (~9K chunks in a 6M-splat 4K first pass):
// a 2048 grid costs ~30 us per frame against 8192, and 16384 gains nothing.
uint32_t mbs_grid(uint32_t max_chunks) { return std::max<uint32_t>(1, std::min<uint32_t>(max_chunks, 8192)); }

}  // namespace

bool span_binning_supported(uint32_t tiles_x, uint32_t tiles_y) {
  return tiles_x >= 1 && tiles_y >= 1 && tiles_x <= kMaxBinTilesX && tiles_y <= kMaxBinTilesY;
}

void span_bin_rects(hipStream_t s, const SpanBinArgs &a, BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals,
                    uint32_t *d_small) {
  const uint32_t groups = div_up(a.cols, 16);
  const uint32_t num_vtiles = a.rows * a.cols;
  const uint32_t cap = std::max<uint32_t>(a.pair_cap, 1);
  uint32_t *d_num_spans = d_small;
  uint32_t *d_num_gspans = d_small + 1;
  uint32_t *d_npairs = d_small + 2;
  // level 1: splats -> row spans (at most one span per pair)
  const size_t n1 = std::max<size_t>(a.n1, 1);
  ws.c1.reserve(n1 * 4);
  ws.base1.reserve(n1 * 4);
  ws.spans.reserve((size_t)cap * 8);
  ws.scan_ws.reserve(scan_workspace_bytes(n1));
  if (a.l1_blocks) {
    k_rows_count<<<a.l1_blocks, kB1, (size_t)a.max_view_rows * kW1 * 4, s>>>(a.rects, a.one, a.tab, a.ntab, a.live,
                                                                            a.d_items, ws.c1.as<uint32_t>());
    DL_HIP(hipGetLastError());
  }
  if (a.d_items) {
    // single view with a device item count: rows * ceil(items / 256) counts
    exclusive_scan_u32_dev(s, ws.c1.as<uint32_t>(), ws.base1.as<uint32_t>(), a.l1_blocks ? a.n1 : 0, a.d_nb1,
                           a.one.rows, d_num_spans, ws.scan_ws.get());
  } else {
    exclusive_scan_u32(s, ws.c1.as<uint32_t>(), ws.base1.as<uint32_t>(), a.l1_blocks ? a.n1 : 0, d_num_spans,
                       ws.scan_ws.get());
  }
  if (a.l1_blocks) {
    const size_t lds1 = (size_t)a.max_view_rows * (kW1 * 4 + 4 + kW1 * 2);
    k_rows_scatter<<<a.l1_blocks, kB1, lds1, s>>>(a.rects, a.one, a.tab, a.ntab, a.live, a.d_items,
                                                  ws.base1.as<uint32_t>(), cap, ws.spans.as<uint2>());
    DL_HIP(hipGetLastError());
  }
  ws.row_start.reserve(((size_t)a.rows + 1) * 4);
  // level 2a: row spans -> 16-column groups (segments = rows)
  const uint32_t max_ch_a = div_up(cap, kChunk) + a.rows;
  if (a.rows <= kSmallSegs) {
    ws.chunk_base_a.reserve(((size_t)a.rows + 1) * 4);
    ws.seg_of_a.reserve((size_t)std::max<uint32_t>(max_ch_a, 1) * 4);
    k_row_start_chunks<<<1, 1024, 0, s>>>(ws.base1.as<uint32_t>(), a.l1_blocks ? a.n1 : 0, a.one, a.tab, a.ntab,
                                          a.rows, a.d_items, d_num_spans, cap, ws.row_start.as<uint32_t>(), max_ch_a,
                                          ws.chunk_base_a.as<uint32_t>(), ws.seg_of_a.as<uint32_t>());
    DL_HIP(hipGetLastError());
  } else {
    k_row_start<<<div_up((size_t)a.rows + 1, 256), 256, 0, s>>>(ws.base1.as<uint32_t>(), a.l1_blocks ? a.n1 : 0,
                                                                a.one, a.tab, a.ntab, a.rows, a.d_items, d_num_spans,
                                                                cap, ws.row_start.as<uint32_t>());
    DL_HIP(hipGetLastError());
    build_chunk_table(s, ws.row_start.as<uint32_t>(), a.rows, ws.tmp, ws.chunk_base_a, ws.seg_of_a, ws.scan_ws,
                      max_ch_a);
  }
  MbsGeom ga{ws.seg_of_a.as<uint32_t>(), ws.chunk_base_a.as<uint32_t>(), ws.row_start.as<uint32_t>(), groups,
             a.rows, max_ch_a, cap, a.live};
  const size_t nc_a = (size_t)kK * max_ch_a;
  ws.c2.reserve(nc_a * 4);
  ws.base2.reserve(nc_a * 4);
  ws.gspans.reserve((size_t)cap * 8);
  ws.scan_ws.reserve(scan_workspace_bytes(nc_a));
  k_mbs_count<0><<<mbs_grid(max_ch_a), kMbsThreads, 0, s>>>(ws.spans.as<uint2>(), ga, ws.c2.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32_dev(s, ws.c2.as<uint32_t>(), ws.base2.as<uint32_t>(), nc_a, ws.chunk_base_a.as<uint32_t>() + a.rows,
                         kK, d_num_gspans, ws.scan_ws.get());
  k_mbs_write<0><<<mbs_grid(max_ch_a), kMbsThreads, 0, s>>>(ws.spans.as<uint2>(), ga, ws.base2.as<uint32_t>(),
                                                    ws.gspans.as<uint2>(), nullptr);
  DL_HIP(hipGetLastError());
  const uint32_t nseg_b = a.rows * groups;
  ws.seg_start_b.reserve(((size_t)nseg_b + 1) * 4);
  // level 2b: group spans -> columns (segments = (row, group))
  const uint32_t max_ch_b = div_up(cap, kChunk) + nseg_b;
  if (nseg_b <= kSmallSegs) {
    ws.chunk_base_b.reserve(((size_t)nseg_b + 1) * 4);
    ws.seg_of_b.reserve((size_t)std::max<uint32_t>(max_ch_b, 1) * 4);
    k_offsets_chunks<<<1, 1024, 0, s>>>(ws.base2.as<uint32_t>(), ws.chunk_base_a.as<uint32_t>(), a.rows, max_ch_a,
                                        nseg_b, groups, a.cols, d_num_gspans, cap, ws.seg_start_b.as<uint32_t>(),
                                        max_ch_b, ws.chunk_base_b.as<uint32_t>(), ws.seg_of_b.as<uint32_t>());
    DL_HIP(hipGetLastError());
  } else {
    k_bucket_offsets<<<div_up((size_t)nseg_b + 1, 256), 256, 0, s>>>(
        ws.base2.as<uint32_t>(), ws.chunk_base_a.as<uint32_t>(), a.rows, max_ch_a, nseg_b, groups, a.cols, 0,
        d_num_gspans, cap, ws.seg_start_b.as<uint32_t>(), nullptr);
    DL_HIP(hipGetLastError());
    build_chunk_table(s, ws.seg_start_b.as<uint32_t>(), nseg_b, ws.tmp, ws.chunk_base_b, ws.seg_of_b, ws.scan_ws,
                      max_ch_b);
  }
  MbsGeom gb{ws.seg_of_b.as<uint32_t>(), ws.chunk_base_b.as<uint32_t>(), ws.seg_start_b.as<uint32_t>(), groups,
             nseg_b, max_ch_b, cap, a.live};
  const size_t nc_b = (size_t)kK * max_ch_b;
  ws.c3.reserve(nc_b * 4);
  ws.base3.reserve(nc_b * 4);
  ws.scan_ws.reserve(scan_workspace_bytes(nc_b));
  k_mbs_count<1><<<mbs_grid(max_ch_b), kMbsThreads, 0, s>>>(ws.gspans.as<uint2>(), gb, ws.c3.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32_dev(s, ws.c3.as<uint32_t>(), ws.base3.as<uint32_t>(), nc_b, ws.chunk_base_b.as<uint32_t>() + nseg_b,
                         kK, d_npairs, ws.scan_ws.get());
  k_mbs_write<1><<<mbs_grid(max_ch_b), kMbsThreads, 0, s>>>(ws.gspans.as<uint2>(), gb, ws.base3.as<uint32_t>(), nullptr,
                                                    pvals);
  DL_HIP(hipGetLastError());
  k_bucket_offsets<<<div_up((size_t)num_vtiles + 1, 256), 256, 0, s>>>(
      ws.base3.as<uint32_t>(), ws.chunk_base_b.as<uint32_t>(), nseg_b, max_ch_b, num_vtiles, groups, a.cols, 1,
      d_npairs, cap, d_off, a.entries);
  DL_HIP(hipGetLastError());
}

void batch_rects(hipStream_t s, const float4 *rec, const uint32_t *ids, const L1View *tab, uint32_t ntab,
                 uint32_t blocks, uint4 *rects) {
  if (!blocks) return;
  k_batch_rects<<<blocks, kB1, 0, s>>>(rec, ids, tab, ntab, rects);
  DL_HIP(hipGetLastError());
}

LiveMap open_sat_tall(hipStream_t s, const uint16_t *pre, uint32_t rows, uint32_t cols, uint32_t *sat,
                      uint32_t *csat) {
  LiveMap lm{pre, cols + 1, sat, csat, div_up(cols, 4) + 1};
  k_open_sat_cols<<<cols + 1, 256, 0, s>>>(pre, rows, cols, sat);
  DL_HIP(hipGetLastError());
  k_open_csat<<<div_up((size_t)(div_up(rows, 4) + 1) * lm.cstride, 256), 256, 0, s>>>(sat, rows, cols, csat);
  DL_HIP(hipGetLastError());
  return lm;
}

void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              const LiveMap &live, const uint2 *rect8, uint32_t n_all, BinWorkspace &ws, uint32_t *d_off,
              uint32_t *pvals, uint32_t *d_small, unsigned long long *d_entries, bool rects_ready) {
  // with rects_ready the caller has written ws.rects (reserved at this size)
  ws.rects.reserve((size_t)std::max<uint32_t>(cn, 1) * 16);
  const bool compact = live.pre && rect8 && cn > 0;
  const uint32_t nblk = div_up(cn, 256);
  if (compact) {
    // live splats only, in depth order
    ws.live_bits.reserve(((size_t)div_up(n_all, 64) + 1) * 8);
    ws.live_cnt.reserve((size_t)nblk * 4);
    ws.live_small.reserve(256);
    uint32_t *ls = ws.live_small.as<uint32_t>();
    ws.live_sat.reserve(((size_t)g.tiles_y + 1) * (g.tiles_x + 1) * 4);
    const uint32_t cstride = div_up(g.tiles_x, 4) + 1, crows = div_up(g.tiles_y, 4) + 1;
    ws.live_csat.reserve((size_t)crows * cstride * 4);
    k_open_sat<<<1, 256, 0, s>>>(live.pre, g.tiles_y, g.tiles_x, ws.live_sat.as<uint32_t>(),
                                 ws.live_csat.as<uint32_t>());
    DL_HIP(hipGetLastError());
    LiveMap lm = live;
    lm.sat = ws.live_sat.as<uint32_t>();
    lm.csat = ws.live_csat.as<uint32_t>();
    lm.cstride = cstride;
    // live bits of every splat (a sequential read of the rects) when the
    // pass holds more than an eighth of them; a smaller pass tests its own
    // splats' rects as it compacts them (two gathers of cn rects)
    unsigned long long *bits = nullptr;
    if ((uint64_t)cn * 8 > n_all) {
      bits = ws.live_bits.as<unsigned long long>();
      k_live_bits<<<div_up(n_all, 256 * kLiveItems), 256, 0, s>>>(rect8, n_all, lm, bits);
      DL_HIP(hipGetLastError());
    }
    const uint32_t grid = div_up(nblk, kLiveItems);
    k_live_count<<<grid, 256, 0, s>>>(ids, cn, bits, rect8, lm, ws.live_cnt.as<uint32_t>());
    DL_HIP(hipGetLastError());
    k_rects_live<<<grid, 256, 0, s>>>(rect8, ids, cn, bits, lm, ws.live_cnt.as<uint32_t>(), ls,
                                      ws.rects.as<uint4>());
    DL_HIP(hipGetLastError());
  } else if (cn && !rects_ready) {
    k_rects<<<nblk, 256, 0, s>>>(rec, ids, cn, g, ws.rects.as<uint4>());
    DL_HIP(hipGetLastError());
  }
  SpanBinArgs a;
  a.rects = ws.rects.as<uint4>();
  a.one = L1View{0, cn, 0, 0, g.tiles_y, 0, g.tiles_x, g.max_cov, cn, 0, 0};
  a.tab = nullptr;
  a.ntab = 0;
  a.l1_blocks = div_up(cn, kB1);
  a.n1 = (size_t)g.tiles_y * a.l1_blocks;
  a.rows = g.tiles_y;
  a.max_view_rows = g.tiles_y;
  a.cols = g.tiles_x;
  a.pair_cap = npairs;
  a.live = live;
  if (compact) {
    a.d_items = ws.live_small.as<uint32_t>();
    a.d_nb1 = ws.live_small.as<uint32_t>() + 1;
  }
  a.entries = d_entries;
  span_bin_rects(s, a, ws, d_off, pvals, d_small);
}

}  // namespace dl
