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

constexpr int kB1 = 256;                 // splats per level-1 block (one per thread)
constexpr int kW1 = kB1 / 32;            // mask words per row
constexpr int kChunk = 1024;             // items per MBS chunk: 4 waves x 4 steps x 64 lanes
constexpr int kK = 16;                   // buckets per MBS segment

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

// row_start[r] = spans before row r (the (row r, block 0) entry of the
// level-1 scan), row_start[tiles_y] = all spans
__global__ __launch_bounds__(256) void k_row_start(const uint32_t *__restrict__ base1, uint32_t nb1,
                                                   uint32_t tiles_y, const uint32_t *__restrict__ d_num_spans,
                                                   uint32_t *__restrict__ row_start) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r < tiles_y) row_start[r] = base1[(size_t)r * nb1];
  if (r == tiles_y) row_start[r] = *d_num_spans;
}

// ----------------------- MBS chunk tables ---------------------------------
__global__ __launch_bounds__(256) void k_seg_chunks(const uint32_t *__restrict__ seg_start, uint32_t nseg,
                                                    uint32_t *__restrict__ nchunks) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s < nseg) nchunks[s] = (seg_start[s + 1] - seg_start[s] + kChunk - 1) / kChunk;
}

__global__ __launch_bounds__(256) void k_seg_of(const uint32_t *__restrict__ chunk_base, uint32_t nseg,
                                                uint32_t *__restrict__ seg_of) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nseg) return;
  for (uint32_t c = chunk_base[s]; c < chunk_base[s + 1]; c++) seg_of[c] = s;
}

// ---------------------------- MBS -----------------------------------------
struct MbsGeom {
  const uint32_t *seg_of;      // chunk -> segment
  const uint32_t *chunk_base;  // segment -> first chunk (nseg + 1 entries)
  const uint32_t *seg_start;   // segment -> first item (nseg + 1 entries)
  uint32_t groups;             // 16-column groups per row (level 2b: segment = row * groups + group)
};

// buckets [lo, hi] of an item; lo > hi: none
template <int LEVEL>
__device__ inline void item_buckets(uint2 it, uint32_t seg, uint32_t groups, uint32_t &lo, uint32_t &hi) {
  const uint32_t xa = it.y & 0xFFFFu, xb = it.y >> 16;
  if (LEVEL == 0) {
    lo = xa >> 4;
    hi = xb >> 4;
  } else {
    const uint32_t g0 = (seg % groups) * 16u;
    lo = max(xa, g0) - g0;
    hi = min(xb, g0 + 15u) - g0;
  }
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
template <int LEVEL>
__global__ __launch_bounds__(256) void k_mbs_count(const uint2 *__restrict__ items, MbsGeom G,
                                                   uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_cnt[4][kK];
  const ChunkSpan cs = chunk_span(G, blockIdx.x);
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t cnt[kK];
#pragma unroll
  for (int b = 0; b < kK; b++) cnt[b] = 0;
#pragma unroll
  for (int step = 0; step < 4; step++) {
    const uint32_t i = cs.i0 + (uint32_t)w * 256u + (uint32_t)step * 64u + (uint32_t)lane;
    uint32_t lo = 1, hi = 0;
    if (i < cs.iend) item_buckets<LEVEL>(items[i], cs.seg, G.groups, lo, hi);
#pragma unroll
    for (int b = 0; b < kK; b++) cnt[b] += (uint32_t)__popcll(__ballot(lo <= (uint32_t)b && (uint32_t)b <= hi));
  }
  if (lane == 0) {
#pragma unroll
    for (int b = 0; b < kK; b++) s_cnt[w][b] = cnt[b];
  }
  __syncthreads();
  if (threadIdx.x < kK) {
    const uint32_t b = threadIdx.x;
    counts[(size_t)kK * cs.cb + (size_t)b * cs.nch + cs.k] = s_cnt[0][b] + s_cnt[1][b] + s_cnt[2][b] + s_cnt[3][b];
  }
}

template <int LEVEL>
__global__ __launch_bounds__(256) void k_mbs_write(const uint2 *__restrict__ items, MbsGeom G,
                                                   const uint32_t *__restrict__ base, uint2 *__restrict__ out_items,
                                                   uint32_t *__restrict__ out_ids) {
  __shared__ uint32_t s_cnt[4][kK];
  const ChunkSpan cs = chunk_span(G, blockIdx.x);
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint2 it[4];
  uint32_t lo[4], hi[4];
  uint32_t cnt[kK];
#pragma unroll
  for (int b = 0; b < kK; b++) cnt[b] = 0;
#pragma unroll
  for (int step = 0; step < 4; step++) {
    const uint32_t i = cs.i0 + (uint32_t)w * 256u + (uint32_t)step * 64u + (uint32_t)lane;
    lo[step] = 1;
    hi[step] = 0;
    it[step] = make_uint2(0u, 0u);
    if (i < cs.iend) {
      it[step] = items[i];
      item_buckets<LEVEL>(it[step], cs.seg, G.groups, lo[step], hi[step]);
    }
#pragma unroll
    for (int b = 0; b < kK; b++)
      cnt[b] += (uint32_t)__popcll(__ballot(lo[step] <= (uint32_t)b && (uint32_t)b <= hi[step]));
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
  for (int step = 0; step < 4; step++) {
#pragma unroll
    for (int b = 0; b < kK; b++) {
      const bool in = lo[step] <= (uint32_t)b && (uint32_t)b <= hi[step];
      const uint64_t m = __ballot(in);
      if (in) {
        const uint32_t pos = run[b] + (uint32_t)__popcll(m & lt);
        if (LEVEL == 0)
          out_items[pos] = it[step];
        else
          out_ids[pos] = it[step].x;
      }
      run[b] += (uint32_t)__popcll(m);
    }
  }
}

// Start of every (segment, bucket) list from the scanned MBS counts: the
// (segment, bucket, chunk 0) entry; a segment without chunks takes the entry
// where it would start, or the total at the end. Level 2a: o = row*groups +
// group. Level 2b: o = tile = row*tiles_x + x -> (row*groups + x/16, x%16).
__global__ __launch_bounds__(256) void k_bucket_offsets(const uint32_t *__restrict__ base,
                                                        const uint32_t *__restrict__ chunk_base, uint32_t n_out,
                                                        uint32_t groups, uint32_t tiles_x, int tile_mode,
                                                        size_t ncounts, const uint32_t *__restrict__ d_total,
                                                        uint32_t *__restrict__ out) {
  const uint32_t o = blockIdx.x * 256 + threadIdx.x;
  if (o > n_out) return;
  if (o == n_out) {
    out[o] = *d_total;
    return;
  }
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
  out[o] = pos < ncounts ? base[pos] : *d_total;
}

// Chunk table of one MBS level: nchunks per segment, its scan (chunk_base,
// with the total in chunk_base[nseg]) and chunk -> segment. Returns the
// number of chunks (host sync).
uint32_t build_chunk_table(hipStream_t s, const uint32_t *seg_start, uint32_t nseg, DeviceBuffer &tmp,
                           DeviceBuffer &chunk_base, DeviceBuffer &seg_of, DeviceBuffer &scan_ws,
                           uint32_t max_chunks) {
  tmp.reserve((size_t)nseg * 4 + 4);
  chunk_base.reserve(((size_t)nseg + 1) * 4);
  seg_of.reserve((size_t)std::max<uint32_t>(max_chunks, 1) * 4);
  scan_ws.reserve(scan_workspace_bytes(nseg));
  k_seg_chunks<<<div_up(nseg, 256), 256, 0, s>>>(seg_start, nseg, tmp.as<uint32_t>());
  DL_HIP(hipGetLastError());
  exclusive_scan_u32(s, tmp.as<uint32_t>(), chunk_base.as<uint32_t>(), nseg, chunk_base.as<uint32_t>() + nseg,
                     scan_ws.get());
  uint32_t total = 0;
  DL_HIP(hipMemcpyAsync(&total, chunk_base.as<uint32_t>() + nseg, 4, hipMemcpyDeviceToHost, s));
  DL_HIP(hipStreamSynchronize(s));
  if (total > max_chunks) fail(DL_ERR_INTERNAL, "span binning: chunk table overflow");
  k_seg_of<<<div_up(nseg, 256), 256, 0, s>>>(chunk_base.as<uint32_t>(), nseg, seg_of.as<uint32_t>());
  DL_HIP(hipGetLastError());
  return total;
}

}  // namespace

bool span_binning_supported(uint32_t tiles_x, uint32_t tiles_y) {
  return tiles_x >= 1 && tiles_y >= 1 && tiles_x <= kMaxBinTilesX && tiles_y <= kMaxBinTilesY;
}

void span_bin(hipStream_t s, const float4 *rec, const uint32_t *ids, uint32_t cn, const BinGrid &g, uint32_t npairs,
              BinWorkspace &ws, uint32_t *d_off, uint32_t *pvals, uint32_t *d_small) {
  const uint32_t num_tiles = g.tiles_x * g.tiles_y;
  const uint32_t groups = div_up(g.tiles_x, 16);
  uint32_t *d_num_spans = d_small;
  uint32_t *d_num_gspans = d_small + 1;
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
  ws.row_start.reserve(((size_t)g.tiles_y + 1) * 4);
  k_row_start<<<div_up((size_t)g.tiles_y + 1, 256), 256, 0, s>>>(ws.base1.as<uint32_t>(), nb1, g.tiles_y, d_num_spans,
                                                                ws.row_start.as<uint32_t>());
  DL_HIP(hipGetLastError());

  // level 2a: row spans -> 16-column groups (segments = rows)
  const uint32_t max_ch_a = div_up(npairs, kChunk) + g.tiles_y;
  const uint32_t nch_a = build_chunk_table(s, ws.row_start.as<uint32_t>(), g.tiles_y, ws.tmp, ws.chunk_base_a,
                                           ws.seg_of_a, ws.scan_ws, max_ch_a);
  MbsGeom ga{ws.seg_of_a.as<uint32_t>(), ws.chunk_base_a.as<uint32_t>(), ws.row_start.as<uint32_t>(), groups};
  const size_t nc_a = (size_t)kK * nch_a;
  ws.c2.reserve(std::max<size_t>(nc_a, 1) * 4);
  ws.base2.reserve(std::max<size_t>(nc_a, 1) * 4);
  ws.gspans.reserve((size_t)std::max<uint32_t>(npairs, 1) * 8);
  ws.scan_ws.reserve(scan_workspace_bytes(std::max<size_t>(nc_a, 1)));
  if (nch_a) {
    k_mbs_count<0><<<nch_a, 256, 0, s>>>(ws.spans.as<uint2>(), ga, ws.c2.as<uint32_t>());
    DL_HIP(hipGetLastError());
  }
  exclusive_scan_u32(s, ws.c2.as<uint32_t>(), ws.base2.as<uint32_t>(), nc_a, d_num_gspans, ws.scan_ws.get());
  if (nch_a) {
    k_mbs_write<0><<<nch_a, 256, 0, s>>>(ws.spans.as<uint2>(), ga, ws.base2.as<uint32_t>(), ws.gspans.as<uint2>(),
                                         nullptr);
    DL_HIP(hipGetLastError());
  }
  const uint32_t nseg_b = g.tiles_y * groups;
  ws.seg_start_b.reserve(((size_t)nseg_b + 1) * 4);
  k_bucket_offsets<<<div_up((size_t)nseg_b + 1, 256), 256, 0, s>>>(ws.base2.as<uint32_t>(), ws.chunk_base_a.as<uint32_t>(),
                                                                   nseg_b, groups, g.tiles_x, 0, nc_a, d_num_gspans,
                                                                   ws.seg_start_b.as<uint32_t>());
  DL_HIP(hipGetLastError());

  // level 2b: group spans -> columns (segments = (row, group))
  const uint32_t max_ch_b = div_up(npairs, kChunk) + nseg_b;
  const uint32_t nch_b = build_chunk_table(s, ws.seg_start_b.as<uint32_t>(), nseg_b, ws.tmp, ws.chunk_base_b,
                                           ws.seg_of_b, ws.scan_ws, max_ch_b);
  MbsGeom gb{ws.seg_of_b.as<uint32_t>(), ws.chunk_base_b.as<uint32_t>(), ws.seg_start_b.as<uint32_t>(), groups};
  const size_t nc_b = (size_t)kK * nch_b;
  ws.c3.reserve(std::max<size_t>(nc_b, 1) * 4);
  ws.base3.reserve(std::max<size_t>(nc_b, 1) * 4);
  ws.scan_ws.reserve(scan_workspace_bytes(std::max<size_t>(nc_b, 1)));
  if (nch_b) {
    k_mbs_count<1><<<nch_b, 256, 0, s>>>(ws.gspans.as<uint2>(), gb, ws.c3.as<uint32_t>());
    DL_HIP(hipGetLastError());
  }
  exclusive_scan_u32(s, ws.c3.as<uint32_t>(), ws.base3.as<uint32_t>(), nc_b, d_npairs, ws.scan_ws.get());
  if (nch_b) {
    k_mbs_write<1><<<nch_b, 256, 0, s>>>(ws.gspans.as<uint2>(), gb, ws.base3.as<uint32_t>(), nullptr, pvals);
    DL_HIP(hipGetLastError());
  }
  k_bucket_offsets<<<div_up((size_t)num_tiles + 1, 256), 256, 0, s>>>(ws.base3.as<uint32_t>(), ws.chunk_base_b.as<uint32_t>(),
                                                                      num_tiles, groups, g.tiles_x, 1, nc_b, d_npairs,
                                                                      d_off);
  DL_HIP(hipGetLastError());
}

}  // namespace dl
