// Forward splat renderer on gfx950: the device path behind dl_render.
//
// Stages (reference citations are relative to the droplet repository):
//   k_cull_project  one coalesced pass over the scene in splat order: the
//                   f64 near/far (or radius) cull and sortable f32 depth key
//                   (src/render/camera.ts:56-99, src/spatial/radix-sort.ts:20-24)
//                   and, for visible splats, the EWA projection, SH colour and
//                   tile count (src/render/shaders/project.ts:22-332)
//   depth sort      stable LSD radix over (key, splat id) (radix-sort.ts:65-117)
//   binning         per-tile lists of splat ids in depth order: span_bin
//                   (binning.hip), or k_emit + stable radix sort + ranges
//                   (emit-pairs.ts:27-56, pipeline.ts:544-578, find-boundaries.ts:17-99)
//   k_rasterize     front-to-back compositing + finalize (rasterize.ts:36-82,
//                   finalize.ts:29-42)
// Projected records stay in splat-id order; the per-tile lists carry splat
// ids, whose order inside a tile is the depth order. The reference's pair
// values (chunk-local depth ranks) are the inverse permutation of these ids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "binning.h"
#include "dl_common.h"
#include "dl_internal.h"
#include "render.h"
#include "render_device.h"

namespace dl {

namespace {

constexpr int kProjBlock = 256;

// One thread per splat in id order, so every scene array is read coalesced.
// Per-block partial sums of visible splats and tile counts replace any
// same-address atomics.
// Per-block partials: visible splats, tile pairs, the least and greatest
// depth key of the visible splats (0xFFFFFFFF / 0 when none), and whether a
// splat's 16-px tile range exceeds maxCoverage (the reference truncates its
// pairs, so only 16-px tiles give its image).
struct CullParts {
  uint32_t *vis;
  unsigned long long *pairs;
  uint32_t *kmin, *kmax, *cov;
};

template <int BANDS, bool COLOR>
__global__ __launch_bounds__(kProjBlock) void k_cull_project(SceneView sc, ProjUniforms u, double near_plane,
                                                             double far_plane, uint32_t *__restrict__ keys,
                                                             uint32_t *__restrict__ ids, float4 *__restrict__ rec,
                                                             uint32_t *__restrict__ counts,
                                                             uint2 *__restrict__ rect8, CullParts parts) {
  __shared__ uint32_t s_vis[kProjBlock / 64], s_kmin[kProjBlock / 64], s_kmax[kProjBlock / 64];
  __shared__ uint32_t s_cov[kProjBlock / 64];
  __shared__ unsigned long long s_pairs[kProjBlock / 64];
  const uint32_t i = blockIdx.x * kProjBlock + threadIdx.x;
  bool vis = false, cov16 = false;
  uint32_t count = 0;
  uint32_t kmin = 0xFFFFFFFFu, kmax = 0u;
  if (i < sc.count) {
    cull_project_one<BANDS, COLOR>(sc, u, near_plane, far_plane, i, i, i, keys, ids, rec, counts, vis, count,
                                   cov16);
    if (vis) kmin = kmax = keys[i];
    // packed tile rect for the live-splat compaction of span binning
    if (rect8) rect8[i] = pack_rect8(splat_rect(rec[3 * (size_t)i], u.tg, u.max_cov));
  }
  uint32_t v = vis ? 1u : 0u;
  unsigned long long pc = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    v += __shfl_xor(v, d, 64);
    pc += __shfl_xor(pc, d, 64);
    kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, d, 64));
    kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, d, 64));
  }
  const bool any_cov = __ballot(cov16) != 0;
  if (__lane_id() == 0) {
    s_vis[threadIdx.x >> 6] = v;
    s_pairs[threadIdx.x >> 6] = pc;
    s_kmin[threadIdx.x >> 6] = kmin;
    s_kmax[threadIdx.x >> 6] = kmax;
    s_cov[threadIdx.x >> 6] = any_cov ? 1u : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tv = 0, tmin = 0xFFFFFFFFu, tmax = 0u, tcov = 0u;
    unsigned long long tp = 0;
#pragma unroll
    for (int w = 0; w < kProjBlock / 64; w++) {
      tv += s_vis[w];
      tp += s_pairs[w];
      tmin = min(tmin, s_kmin[w]);
      tmax = max(tmax, s_kmax[w]);
      tcov |= s_cov[w];
    }
    parts.vis[blockIdx.x] = tv;
    parts.pairs[blockIdx.x] = tp;
    parts.kmin[blockIdx.x] = tmin;
    parts.kmax[blockIdx.x] = tmax;
    parts.cov[blockIdx.x] = tcov;
  }
}

// Frame totals from the cull's partials: d_nvis[0] visible splats,
// d_nvis[2] / d_nvis[3] the least / greatest visible depth key, d_nvis[7]
// whether maxCoverage truncates a splat at 16-px tiles, *d_pairs the tile
// pairs. kReduceGroups workgroups each reduce a slice of the partials and
// fold it into the totals with atomics (integer sums, min, max and or, so
// the order does not matter); k_frame_init set the totals' start values. One
// workgroup reading all 23K partials of a 6M-splat frame took 12 us.
constexpr uint32_t kReduceGroups = 32;

__global__ __launch_bounds__(1024) void k_reduce_totals(CullParts parts, uint32_t nparts,
                                                        uint32_t *__restrict__ d_nvis,
                                                        unsigned long long *__restrict__ d_pairs) {
  __shared__ uint32_t s_v[16], s_lo[16], s_hi[16], s_c[16];
  __shared__ unsigned long long s_p[16];
  uint32_t v = 0, lo = 0xFFFFFFFFu, hi = 0u, c = 0u;
  unsigned long long p = 0;
  const uint32_t per = (nparts + gridDim.x - 1) / gridDim.x;
  const uint32_t k0 = blockIdx.x * per, k1 = min(nparts, k0 + per);
  // independent iterations, unrolled so that their loads are in flight together
#pragma unroll 4
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 1024) {
    v += parts.vis[k];
    p += parts.pairs[k];
    lo = min(lo, parts.kmin[k]);
    hi = max(hi, parts.kmax[k]);
    c |= parts.cov[k];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    v += __shfl_xor(v, d, 64);
    p += __shfl_xor(p, d, 64);
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, d, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, d, 64));
    c |= (uint32_t)__shfl_xor((int)c, d, 64);
  }
  if (__lane_id() == 0) {
    s_v[threadIdx.x >> 6] = v;
    s_p[threadIdx.x >> 6] = p;
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
    s_c[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tv = 0, tlo = 0xFFFFFFFFu, thi = 0u, tc = 0u;
    unsigned long long tp = 0;
    for (int w = 0; w < 16; w++) {
      tv += s_v[w];
      tp += s_p[w];
      tlo = min(tlo, s_lo[w]);
      thi = max(thi, s_hi[w]);
      tc |= s_c[w];
    }
    if (k0 < k1) {
      atomicAdd(&d_nvis[0], tv);
      atomicMin(&d_nvis[2], tlo);
      atomicMax(&d_nvis[3], thi);
      atomicOr(&d_nvis[7], tc);
      atomicAdd(d_pairs, tp);
    }
  }
}

__global__ __launch_bounds__(256) void k_gather_counts(const uint32_t *__restrict__ counts,
                                                       const uint32_t *__restrict__ ids, uint32_t cn,
                                                       uint32_t *__restrict__ out) {
  const uint32_t jj = blockIdx.x * 256 + threadIdx.x;
  if (jj < cn) out[jj] = counts[ids[jj]];
}

// ---------------------------------------------------------------------------
// Pass cuts of a frame above the pair budget, on the device: P[j] = the tile
// counts of the first j splats in depth order (u64, P[0] = 0), then one wave
// cuts each pass greedily by 64-way searches on P. A host loop over the n
// counts gives the same cuts after copying them (24 MB at 6M splats) and
// walking them on one core, which left the GPU idle for milliseconds.
// ---------------------------------------------------------------------------
constexpr uint32_t kCutBlock = 1024;   // splats per block of the prefix kernels (256 threads x 4)
constexpr uint32_t kMaxCutSegs = 128;  // passes before the cut (more: the host loop)

struct CutSegs {
  uint32_t n;                      // passes to cut
  uint32_t at[kMaxCutSegs + 1];    // their starts, then their end
};

// g[j] = counts[ids[j]] and the block's sum (kCutBlock splats a block)
__global__ __launch_bounds__(256) void k_cut_gather(const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ ids, uint32_t n,
                                                    uint32_t *__restrict__ g, unsigned long long *__restrict__ bsum) {
  __shared__ unsigned long long s_w[4];
  const uint32_t base = blockIdx.x * kCutBlock;
  unsigned long long sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    const uint32_t j = base + i * 256 + threadIdx.x;
    if (j < n) {
      const uint32_t c = counts[ids[j]];
      g[j] = c;
      sum += c;
    }
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__device__ inline unsigned long long wave_inclusive_scan_u64(unsigned long long v, uint32_t lane) {
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan of the nb block sums in place (one workgroup)
__global__ __launch_bounds__(1024) void k_cut_scan_blocks(unsigned long long *__restrict__ bsum, uint32_t nb) {
  __shared__ unsigned long long s_w[16];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += 1024) {
    const uint32_t j = c0 + threadIdx.x;
    const unsigned long long v = j < nb ? bsum[j] : 0ull;
    const unsigned long long inc = wave_inclusive_scan_u64(v, lane);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    unsigned long long wbase = 0, tot = 0;
    for (uint32_t k = 0; k < 16; k++) {
      if (k < w) wbase += s_w[k];
      tot += s_w[k];
    }
    if (j < nb) bsum[j] = carry + wbase + inc - v;
    carry += tot;
    __syncthreads();
  }
}

// P[j + 1] = bsum[block] + the counts of the block up to j; P[0] = 0
__global__ __launch_bounds__(256) void k_cut_prefix(const uint32_t *__restrict__ g,
                                                    const unsigned long long *__restrict__ bsum, uint32_t n,
                                                    unsigned long long *__restrict__ P) {
  __shared__ unsigned long long s_w[4];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t j0 = blockIdx.x * kCutBlock + threadIdx.x * 4;
  uint32_t c[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) c[i] = j0 + i < n ? g[j0 + i] : 0u;
  const unsigned long long t = (unsigned long long)c[0] + c[1] + c[2] + c[3];
  const unsigned long long inc = wave_inclusive_scan_u64(t, lane);
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  unsigned long long run = bsum[blockIdx.x] + inc - t;
  for (uint32_t k = 0; k < w; k++) run += s_w[k];
  if (blockIdx.x == 0 && threadIdx.x == 0) P[0] = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    run += c[i];
    if (j0 + i < n) P[j0 + i + 1] = run;
  }
}

// One wave: each pass [a, e) of segs is cut at b = the first index past a
// whose prefix from a exceeds cap, less one, and at least a + 1 (a splat
// alone above cap is a pass of its own): out[0] = the number of passes
// (~0 when more than maxout), out[1 + 2i] = the end of pass i, out[2 + 2i]
// = its pair count P[b] - P[a].
__global__ __launch_bounds__(64) void k_cut_greedy(const unsigned long long *__restrict__ P, CutSegs segs,
                                                   unsigned long long cap, unsigned long long *__restrict__ out,
                                                   uint32_t maxout) {
  const uint32_t lane = threadIdx.x;
  uint32_t cnt = 0;
  bool over = false;
  for (uint32_t sg = 0; sg < segs.n; sg++) {
    uint32_t a = segs.at[sg];
    const uint32_t e = segs.at[sg + 1];
    while (a < e) {
      const unsigned long long pa = P[a];
      const unsigned long long target = pa + cap;
      // the first k in [a + 1, e] with P[k] > target, e + 1 when none: P is
      // non-decreasing, so the probes above target are a suffix of the lanes
      uint64_t lo = (uint64_t)a + 1, hi = (uint64_t)e + 1;
      while (lo < hi) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t k = lo + lane * step;
        const bool gt = k < hi && P[k] > target;
        const unsigned long long m = __ballot(gt);
        if (m == 0) {
          const uint64_t last = lo + std::min<uint64_t>(63, (hi - 1 - lo) / step) * step;
          lo = last + 1;
        } else {
          const uint32_t f = (uint32_t)__builtin_ctzll(m);
          hi = lo + f * step;
          if (f > 0) lo = lo + (f - 1) * step + 1;
        }
      }
      const uint32_t b = lo == (uint64_t)e + 1 ? e : (uint32_t)std::max<uint64_t>((uint64_t)a + 1, lo - 1);
      if (cnt < maxout) {
        if (lane == 0) {
          out[1 + 2 * cnt] = b;
          out[2 + 2 * cnt] = P[b] - pa;
        }
        cnt++;
      } else {
        over = true;
      }
      a = b;
    }
  }
  if (lane == 0) out[0] = over ? ~0ull : (unsigned long long)cnt;
}

// ---------------------------------------------------------------------------
// emit (emit-pairs.ts:27-56); values are splat ids
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_emit(const float4 *__restrict__ rec, const uint32_t *__restrict__ ids,
                                              const uint32_t *__restrict__ offs, uint32_t cn, TileGrid tg,
                                              uint32_t max_cov, uint32_t *__restrict__ pkeys,
                                              uint32_t *__restrict__ pvals) {
  const uint32_t jj = blockIdx.x * 128 + threadIdx.x;
  if (jj >= cn) return;
  const uint32_t id = ids[jj];
  const float4 v0 = rec[3 * (size_t)id];
  const float sx = v0.x, sy = v0.y, radius = v0.z;
  if (radius <= 0.0f) return;
  uint32_t min_tx, max_tx, min_ty, max_ty;
  if (!tile_range(sx, sy, radius, tg, min_tx, max_tx, min_ty, max_ty)) return;
  const uint32_t tiles_x = tg.tiles_x;
  uint32_t off = offs[jj];
  uint32_t written = 0;
  for (uint32_t ty = min_ty; ty <= max_ty; ty++) {
    for (uint32_t tx = min_tx; tx <= max_tx; tx++) {
      if (written >= max_cov) return;
      pkeys[off] = ty * tiles_x + tx;
      pvals[off] = id;
      off++;
      written++;
    }
  }
}

// Six waves per SIMD: the staging loop keeps the tile's accumulators in LDS,
// so the kernel fits 80 VGPRs without spilling, and the extra waves hide the
// latency of the record loads (measured 0.73 -> 0.72 ms at 6M splats / 4K).
template <bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_rasterize(const float4 *__restrict__ rec, const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ pvals, float4 *__restrict__ accum,
                                                  uint32_t *__restrict__ out_rgba, RastParams p,
                                                  unsigned long long *__restrict__ d_iters) {
  __shared__ RastLds lds;
  const uint32_t tile = p.order ? p.order[blockIdx.x] : blockIdx.x;
  const uint32_t start = p.has_pairs ? off[tile] : 0u, end = p.has_pairs ? off[tile + 1] : 0u;
  rasterize_tile<COUNT>(tile % p.tiles_x, tile / p.tiles_x, tile, start, end, rec, pvals, accum, out_rgba, p, d_iters,
                        lds);
}

// The autotuner's other rasterizer shapes (single-view frames without
// counting): tiles of 8 or 16 px in 2x2 groups per 256-thread workgroup, 8-px
// tiles one wave each, and 32-px tile lists rasterized one 16-px quadrant
// per wave. A group or a quadrant is one launch unit: p.order lists units
// and p.cost_key records one key per unit.
template <uint32_t TILE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_rasterize_2x2(
    const float4 *__restrict__ rec, const uint32_t *__restrict__ off, const uint32_t *__restrict__ pvals,
    float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba, RastParams p) {
  __shared__ RastLds lds[4];
  __shared__ uint32_t s_work[4];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t unit = p.order ? p.order[blockIdx.x] : blockIdx.x;
  const uint32_t ux = (p.tiles_x + 1) / 2;
  const uint32_t tx = 2 * (unit % ux) + (w & 1), ty = 2 * (unit / ux) + (w >> 1);
  uint32_t work = 0;
  if (tx < p.tiles_x && ty < p.tiles_y) {
    const uint32_t tile = ty * p.tiles_x + tx;
    const uint32_t start = p.has_pairs ? off[tile] : 0u, end = p.has_pairs ? off[tile + 1] : 0u;
    if constexpr (TILE == 16)
      work = rasterize_tile16<false, true>(tx, ty, tile, start, end, rec, pvals, accum, out_rgba, p, nullptr, lds[w]);
    else
      work = rasterize_tile8<true>(tx, ty, tile, start, end, rec, pvals, accum, out_rgba, p, lds[w]);
  }
  if (p.cost_key) {
    if ((threadIdx.x & 63) == 0) s_work[w] = work;
    __syncthreads();
    if (threadIdx.x == 0) p.cost_key[unit] = work_key(s_work[0] + s_work[1] + s_work[2] + s_work[3]);
  }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_rasterize8(
    const float4 *__restrict__ rec, const uint32_t *__restrict__ off, const uint32_t *__restrict__ pvals,
    float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba, RastParams p) {
  __shared__ RastLds lds;
  const uint32_t tile = p.order ? p.order[blockIdx.x] : blockIdx.x;
  const uint32_t start = p.has_pairs ? off[tile] : 0u, end = p.has_pairs ? off[tile + 1] : 0u;
  const uint32_t work =
      rasterize_tile8<false>(tile % p.tiles_x, tile / p.tiles_x, tile, start, end, rec, pvals, accum, out_rgba, p, lds);
  if (p.cost_key && threadIdx.x == 0) p.cost_key[tile] = work_key(work);
}

// 32-px tile lists, one wave64 workgroup per 16-px quadrant (p.tiles_x is
// the 16-px grid's width): quadrant q reads the list of its 32-px tile
// (p.list_tiles_x wide) and keeps to the splats whose 16-px range holds it.
// Closed flags, work keys and the launch order are per quadrant. Four
// quadrants as the waves of one workgroup measured 680 us against 585 us
// at 16-px tiles on the 6M-splat 4K frame: the workgroup held its slots
// until its slowest quadrant ended.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_rasterize32(
    const float4 *__restrict__ rec, const uint32_t *__restrict__ off, const uint32_t *__restrict__ pvals,
    float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba, RastParams p) {
  __shared__ RastLds lds;
  const uint32_t q = p.order ? p.order[blockIdx.x] : blockIdx.x;
  const uint32_t qx = q % p.tiles_x, qy = q / p.tiles_x;
  const uint32_t tile = (qy >> 1) * p.list_tiles_x + (qx >> 1);
  const uint32_t start = p.has_pairs ? off[tile] : 0u, end = p.has_pairs ? off[tile + 1] : 0u;
  uint32_t work = 0;
  if (!tile_skipped(p, q, start, end)) {
    Px16 s;
    px16_load(s, qx * kTile, qy * kTile, accum, p);
    work = px16_run_quad<false>(s, qx * kTile, qy * kTile, start, end, rec, pvals, p, lds, qx, qy);
    bool closing = false;
    if (p.closed && !p.finalize) {
      const bool open = px16_open(s);
      if (threadIdx.x == 0) p.closed[q] = open ? 0 : 1;
      closing = !open && p.final_on_close;
    }
    px16_store(s, p, accum, out_rgba, p.write_accum && !closing, p.finalize || closing);
  }
  if (p.cost_key && threadIdx.x == 0) p.cost_key[q] = work_key(work);
}

// Open map of the 32-px grid (cols x rows) from the 16-px quadrants' closed
// flags (t16x wide, t16y rows): a 32-px tile is open while a quadrant is.
// Per row, prefix counts of open tiles, as k_open_prefix. One 256-thread
// block per row (cols <= 256).
__global__ __launch_bounds__(256) void k_open_prefix_quads(const uint8_t *__restrict__ closed16, uint32_t t16x,
                                                           uint32_t t16y, uint32_t cols,
                                                           uint16_t *__restrict__ pre) {
  __shared__ uint32_t s_w[4];
  const uint32_t y = blockIdx.x, x = threadIdx.x;
  uint32_t open = 0;
  if (x < cols) {
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      const uint32_t qx = 2 * x + (k & 1), qy = 2 * y + (k >> 1);
      if (qx < t16x && qy < t16y && !closed16[(size_t)qy * t16x + qx]) open = 1;
    }
  }
  uint32_t inc = open;
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_w[x >> 6] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t w = 0; w < (x >> 6); w++) off += s_w[w];
  inc += off;
  uint16_t *r = pre + (size_t)y * (cols + 1);
  if (x == 0) r[0] = 0;
  if (x < cols) r[x + 1] = (uint16_t)inc;
}

// Start of a frame: the list-entry counter is zeroed and the cull totals
// take the start values that k_reduce_totals folds into (visible 0, least
// key ~0, greatest key 0, coverage flag 0, pairs 0).
__global__ void k_frame_init(unsigned long long *__restrict__ entries, uint32_t *__restrict__ d_nvis,
                             unsigned long long *__restrict__ d_pairs) {
  if (threadIdx.x == 0) {
    *entries = 0;
    d_nvis[0] = 0;
    d_nvis[2] = 0xFFFFFFFFu;
    d_nvis[3] = 0;
    d_nvis[7] = 0;
    *d_pairs = 0;
  }
}

__global__ void k_add_const(uint32_t n, unsigned long long *__restrict__ acc) {
  if (threadIdx.x == 0) *acc += n;
}

// Counting frames: the pass's binned splats (live ones in a compacted pass,
// else cn), row spans and group spans (span-binning scalars [0], [1]) added
// to the frame's tallies.
__global__ void k_bin_tally(const uint32_t *__restrict__ bin, const uint32_t *__restrict__ live_items, uint32_t cn,
                            unsigned long long *__restrict__ tally) {
  if (threadIdx.x == 0) {
    tally[0] += live_items ? min(live_items[0], cn) : cn;
    tally[1] += bin[0];
    tally[2] += bin[1];
  }
}


// Rasterizer launch for tiles of `tile` px in groups of group x group tiles
// per workgroup (32-px tiles: one per workgroup); units = workgroups.
void launch_rasterize(bool count, uint32_t tile, uint32_t group, uint32_t units, hipStream_t s, const float4 *rec,
                      const uint32_t *off, const uint32_t *pvals, float4 *accum, uint32_t *out_rgba,
                      const RastParams &p, unsigned long long *d_iters) {
  if (tile == kTile && group == 1) {
    if (count)
      k_rasterize<true><<<units, 64, 0, s>>>(rec, off, pvals, accum, out_rgba, p, d_iters);
    else
      k_rasterize<false><<<units, 64, 0, s>>>(rec, off, pvals, accum, out_rgba, p, d_iters);
  } else {
    if (count) fail(DL_ERR_INTERNAL, "counting frames rasterize 16-px tiles one per workgroup");
    if (tile == 32)
      k_rasterize32<<<units, 64, 0, s>>>(rec, off, pvals, accum, out_rgba, p);
    else if (group == 2 && tile == kTile)
      k_rasterize_2x2<16><<<units, 256, 0, s>>>(rec, off, pvals, accum, out_rgba, p);
    else if (group == 2)
      k_rasterize_2x2<8><<<units, 256, 0, s>>>(rec, off, pvals, accum, out_rgba, p);
    else
      k_rasterize8<<<units, 64, 0, s>>>(rec, off, pvals, accum, out_rgba, p);
  }
  DL_HIP(hipGetLastError());
}

using CullProjFn = void (*)(SceneView, ProjUniforms, double, double, uint32_t *, uint32_t *, float4 *, uint32_t *,
                            uint2 *, CullParts);

CullProjFn cull_project_for(uint32_t bands, bool color) {
  switch (bands) {
    case 0: return color ? k_cull_project<0, true> : k_cull_project<0, false>;
    case 1: return color ? k_cull_project<1, true> : k_cull_project<1, false>;
    case 2: return color ? k_cull_project<2, true> : k_cull_project<2, false>;
    default: return color ? k_cull_project<3, true> : k_cull_project<3, false>;
  }
}

using ColorFn = void (*)(SceneView, const uint32_t *, const uint4 *, uint32_t, const uint32_t *, float4 *);

// The first pass's level-1 records and colours in one gather of each
// splat's projection record: item j names splat ids[j]. Same arithmetic as
// k_rects followed by k_splat_colors.
template <int BANDS>
__global__ __launch_bounds__(256) void k_rects_colors(SceneView sc, const uint32_t *__restrict__ ids, uint32_t cn,
                                                      BinGrid g, float4 *__restrict__ rec,
                                                      uint4 *__restrict__ rects) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= cn) return;
  const uint32_t id = ids[j];
  const float4 v0 = rec[3 * (size_t)id];
  float *slot = reinterpret_cast<float *>(rec + 3 * (size_t)id + 2);
  const float pcx = slot[0], pcy = slot[1], pcz = slot[2];
  rects[j] = pack_rect(id, splat_rect(v0, g.tg, g.max_cov));
  float col[3];
  splat_color<BANDS>(sc, id, pcx, pcy, pcz, col);
  slot[0] = col[0];
  slot[1] = col[1];
  slot[2] = col[2];
}

void launch_rects_colors(hipStream_t s, const SceneView &sc, const uint32_t *ids, uint32_t cn, const BinGrid &g,
                         float4 *rec, uint4 *rects) {
  if (cn == 0) return;
  using Fn = void (*)(SceneView, const uint32_t *, uint32_t, BinGrid, float4 *, uint4 *);
  Fn f = sc.bands == 0 ? k_rects_colors<0> : sc.bands == 1 ? k_rects_colors<1>
       : sc.bands == 2 ? k_rects_colors<2> : k_rects_colors<3>;
  f<<<div_up(cn, 256), 256, 0, s>>>(sc, ids, cn, g, rec, rects);
  DL_HIP(hipGetLastError());
}

void launch_colors(hipStream_t s, const SceneView &sc, const uint32_t *ids, const uint4 *rects, uint32_t n_cap,
                   const uint32_t *d_n, float4 *rec) {
  if (n_cap == 0) return;
  ColorFn f = sc.bands == 0 ? k_splat_colors<0> : sc.bands == 1 ? k_splat_colors<1>
            : sc.bands == 2 ? k_splat_colors<2> : k_splat_colors<3>;
  f<<<div_up(n_cap, 256), 256, 0, s>>>(sc, ids, rects, n_cap, d_n, rec);
  DL_HIP(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Stage events only timestamp the stream: without the system-scope fence a
// record costs no cache writeback and the GPU does not idle at stage
// boundaries (timestamps are read after the frame's own synchronisation).
void Renderer::ensure_events(size_t n) {
  while (events.size() < n) {
    hipEvent_t e;
    DL_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    events.push_back(e);
  }
}

void *Renderer::pinned_small() {
  if (!host_small) DL_HIP(hipHostMalloc(&host_small, kHostSmallBytes, hipHostMallocDefault));
  return host_small;
}

void Renderer::mark(int stage) {
  if (!timing) return;
  ensure_events(ev_used + 1);
  DL_HIP(hipEventRecord(events[ev_used], ctx->stream));
  ev_stage.push_back(stage);
  ev_used++;
}

template <class T>
static void tap_append(std::vector<uint8_t> &dst, const T *src, size_t count) {
  const uint8_t *b = reinterpret_cast<const uint8_t *>(src);
  dst.insert(dst.end(), b, b + count * sizeof(T));
}

template <class T>
static std::vector<T> to_host(const T *dev, size_t count, hipStream_t s) {
  std::vector<T> h(count);
  if (count) DL_HIP(hipMemcpyAsync(h.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, s));
  DL_HIP(hipStreamSynchronize(s));
  return h;
}

// pipeline.ts:475-506: the project uniforms as the f32 values the
// reference writes (focal from the f64 fovY)
ProjUniforms make_uniforms(const dl_camera &cam, uint32_t max_cov, uint32_t tile) {
  ProjUniforms u{};
  std::memcpy(u.view, cam.view_matrix, sizeof(u.view));
  u.width = (float)cam.width;
  u.height = (float)cam.height;
  u.focal = (float)((double)cam.height / (2.0 * std::tan(cam.fov_y / 2.0)));
  u.cx = (float)((double)cam.width / 2.0);
  u.cy = (float)((double)cam.height / 2.0);
  u.near_ = (float)cam.near_plane;
  u.far_ = (float)cam.far_plane;
  u.tg = tile_grid(cam.width, cam.height, tile);
  u.tiles_x = u.tg.tiles_x;
  u.tiles_y = u.tg.tiles_y;
  u.max_cov = max_cov_for_tile(max_cov, tile);
  u.proj = cam.projection;
  u.max_cov16 = max_cov;
  return u;
}

void Renderer::render(const SceneView &sc, const dl_camera &cam, const dl_background &bg) {
  if (cam.width == 0 || cam.height == 0) fail(DL_ERR_INVALID, "camera width and height must be > 0");
  if (sc.bands > 3) fail(DL_ERR_INVALID, "shBands must be 0..3");
  if (cam.projection != DL_PINHOLE && cam.projection != DL_EQUIRECT) fail(DL_ERR_INVALID, "unknown projection");
  // any other render replaces the lists, records and accumulator that a
  // backward would read
  if (!train_mode) train_valid = false;
  hipStream_t s = ctx->stream;
  ctx->make_current();
  ev_used = 0;
  ev_stage.clear();
  for (auto &t : taps) t.clear();

  if (opts.tile_size != 0 && opts.tile_size != 8 && opts.tile_size != 16 && opts.tile_size != 32)
    fail(DL_ERR_INVALID, "tile_size must be 0, 8, 16 or 32");
  if (opts.raster_group > 2) fail(DL_ERR_INVALID, "raster_group must be 0, 1 or 2");

  const uint32_t n = sc.count;
  const uint32_t W = cam.width, H = cam.height;
  // tile size and rasterizer group of this frame: 16 px, one wave per tile
  // where the lists or counts must be the reference's
  uint32_t T = opts.tile_size ? opts.tile_size : kTile;
  uint32_t G = opts.raster_group == 2 ? 2u : 1u;
  if (keep_taps || train_mode || counting || force_tile16 || last_cov16) T = kTile;
  if (keep_taps || train_mode || counting) G = 1;
  if (T == 32) G = 1;
  const TileGrid tg = tile_grid(W, H, T);
  const uint32_t tiles_x = tg.tiles_x, tiles_y = tg.tiles_y;
  const uint32_t num_tiles = tiles_x * tiles_y;
  // workgroups of the rasterizer launch: tiles, 2x2 groups of them, or the
  // 16-px quadrants of 32-px tiles (which also carry the closed flags)
  const uint32_t num16 = tg.t16x * tg.t16y;
  const uint32_t units = G == 2 ? div_up(tiles_x, 2) * div_up(tiles_y, 2) : T == 32 ? num16 : num_tiles;
  const uint32_t flag_tiles = T == 32 ? num16 : num_tiles;
  const uint32_t shape = T * 4 + G;
  const size_t npx = (size_t)W * H;
  stats = dl_render_stats{};
  stats.num_tiles = num_tiles;
  stats.tiles_x = tiles_x;
  stats.tiles_y = tiles_y;
  stats.tile_size = T;
  stats.raster_group = G;
  width = W;
  height = H;

  rgba.reserve(npx * 4);
  small.reserve(256);
  // device scalars: [0] visible, [1] chunk pairs, [2..3] least / greatest
  // visible depth key, [4..6] span-binning scalars, [7] maxCoverage
  // truncation at 16-px tiles, [8..9] u64 total pairs, [10..15] u64 loop iterations and
  // consumed list entries (counting frames), list entries built, [16..21]
  // u64 binned splats, row spans and group spans (counting and tally frames)
  uint32_t *d_nvis = small.as<uint32_t>();
  uint32_t *d_total = d_nvis + 1;
  uint32_t *d_bin = d_nvis + 4;
  unsigned long long *d_pairs_all = reinterpret_cast<unsigned long long *>(d_nvis + 8);
  unsigned long long *d_iters = reinterpret_cast<unsigned long long *>(d_nvis + 10);
  unsigned long long *d_entries = d_iters + 2;
  unsigned long long *d_tally = reinterpret_cast<unsigned long long *>(d_nvis + 16);
  // binning tallies: counting frames and list-tally frames (which keep
  // the frame's tile size and rasterizer)
  const bool tally = counting || tallying;
  if (counting) DL_HIP(hipMemsetAsync(d_iters, 0, 2 * sizeof(unsigned long long), s));
  if (tally) DL_HIP(hipMemsetAsync(d_tally, 0, 3 * sizeof(unsigned long long), s));
  k_frame_init<<<1, 64, 0, s>>>(d_entries, d_nvis, d_pairs_all);
  DL_HIP(hipGetLastError());
  row_iters_rows = 0;
  if (counting) {
    row_iters.reserve((size_t)tiles_y * 8);
    DL_HIP(hipMemsetAsync(row_iters.get(), 0, (size_t)tiles_y * 8, s));
    row_iters_rows = tiles_y;
  }

  RastParams rp{};
  rp.tiles_x = tiles_x;
  rp.width = W;
  rp.height = H;
  rp.bg_r = bg.r;
  rp.bg_g = bg.g;
  rp.bg_b = bg.b;
  rp.bg_a = bg.a;
  rp.row_iters = counting ? row_iters.as<unsigned long long>() : nullptr;
  rp.tiles_y = tiles_y;
  rp.t16x = tg.t16x;
  rp.t16y = tg.t16y;
  rp.list_tiles_x = tiles_x;
  if (T == 32) rp.tiles_x = tg.t16x;  // the quadrant grid

  const ProjUniforms u = make_uniforms(cam, max_cov, T);

  // ---- cull + project (splat order), depth sort ----
  uint32_t nvis = 0;
  unsigned long long pairs_all = 0;
  const uint32_t *sorted_ids = nullptr;
  // packed rects (id order) for the live-splat compaction of later passes
  const bool want_rect8 = opts.binning == 0 && span_binning_supported(tiles_x, tiles_y) && !keep_taps &&
                          !train_mode && !opts.full_lists;
  const bool speculative = !force_exact && spec_cap > 0 && n > 0 && !keep_taps && !train_mode &&
                           opts.binning == 0 && span_binning_supported(tiles_x, tiles_y) &&
                           (opts.chunk_cap == 0 || opts.chunk_cap >= n);
  // A speculative frame that will run in passes leaves the colours out of the
  // projection; each pass computes them for the splats that reach its tile
  // lists (all splats of the first pass, the live ones of later passes).
  bool defer_color = false;
  if (speculative && !opts.full_lists &&
      (opts.first_pass || opts.pass_growth || last_pairs_all >= kPassMinPairs)) {
    const uint64_t len = opts.first_pass ? opts.first_pass : std::max<uint64_t>(kPassMinFirst, n / kPassFirstDiv);
    defer_color = len < n;
  }
  if (n > 0) {
    const uint32_t nparts = div_up(n, kProjBlock);
    keys.reserve((size_t)n * 4);
    ids.reserve((size_t)n * 4);
    tkeys.reserve((size_t)n * 4);
    tids.reserve((size_t)n * 4);
    rec.reserve((size_t)n * 48);
    if (want_rect8) rect8.reserve((size_t)n * 8);
    counts.reserve((size_t)n * 4);
    const size_t pstride = ((size_t)nparts * 8 + 255) & ~(size_t)255;
    parts.reserve(5 * pstride);
    sort_ws.reserve(radix_workspace_bytes(n, 32));
    CullParts cp{};
    cp.pairs = parts.as<unsigned long long>();
    cp.vis = reinterpret_cast<uint32_t *>(parts.as<uint8_t>() + pstride);
    cp.kmin = reinterpret_cast<uint32_t *>(parts.as<uint8_t>() + 2 * pstride);
    cp.kmax = reinterpret_cast<uint32_t *>(parts.as<uint8_t>() + 3 * pstride);
    cp.cov = reinterpret_cast<uint32_t *>(parts.as<uint8_t>() + 4 * pstride);
    mark(DL_STAGE_CULL);
    cull_project_for(sc.bands, !defer_color)<<<nparts, kProjBlock, 0, s>>>(sc, u, cam.near_plane, cam.far_plane,
                                                             keys.as<uint32_t>(),
                                                             speculative ? nullptr : ids.as<uint32_t>(),
                                                             rec.as<float4>(),
                                                             speculative ? nullptr : counts.as<uint32_t>(),
                                                             want_rect8 ? rect8.as<uint2>() : nullptr, cp);
    DL_HIP(hipGetLastError());
    k_reduce_totals<<<std::min(kReduceGroups, nparts), 1024, 0, s>>>(cp, nparts, d_nvis, d_pairs_all);
    DL_HIP(hipGetLastError());
    mark(DL_STAGE_DEPTH_SORT);
    // A speculative frame sorts on key - (least visible key) over the span
    // bits of the last frame (plus key_headroom), with up to 9-bit digits,
    // when that takes fewer passes; invisible splats map past every visible
    // key, so the order is the 32-bit sort's whenever the span fits (checked
    // after the frame).
    int sort_bits = 32, sort_digits = digit_bits();
    RadixKeyMap km{};
    if (speculative && spec_key_bits > 0 && spec_key_bits < 32) {
      const int md = digit_bits() >= 8 ? 9 : digit_bits();
      if (radix_passes((int)spec_key_bits, md) < radix_passes(32, digit_bits())) {
        sort_bits = (int)spec_key_bits;
        sort_digits = md;
        km = RadixKeyMap{d_nvis + 2, kInvisibleKey, (uint32_t)((1ull << spec_key_bits) - 1ull)};
      }
    }
    key_bits_used = km.kmin ? (uint32_t)sort_bits : 0u;
    stats.depth_key_bits = (uint32_t)sort_bits;
    // a speculative frame needs no per-splat tile counts (span binning
    // counts its own) and lets the sort's first pass make the ids 0..n-1
    const bool in_tmp = radix_sort_pairs(s, keys.as<uint32_t>(), ids.as<uint32_t>(), tkeys.as<uint32_t>(),
                                         tids.as<uint32_t>(), n, sort_bits, sort_ws.get(), sort_digits, km,
                                         speculative);
    sorted_ids = in_tmp ? tids.as<uint32_t>() : ids.as<uint32_t>();
    const uint32_t *sorted_keys = in_tmp ? tkeys.as<uint32_t>() : keys.as<uint32_t>();
    if (speculative) {
      // invisible splats sort last and have no tiles: passes over all n
      // splats build the same lists; the capacity is checked after the frame
      nvis = n;
      pairs_all = spec_cap;
    } else {
      // small[0..7] and the pair total after them, in one pinned read
      uint32_t head[10];
      static_assert(sizeof(head) <= kHostSmallBytes, "pinned readback too small");
      void *hs = pinned_small();
      DL_HIP(hipMemcpyAsync(hs, d_nvis, sizeof(head), hipMemcpyDeviceToHost, s));
      DL_HIP(hipStreamSynchronize(s));
      std::memcpy(head, hs, sizeof(head));
      pairs_all = (unsigned long long)head[8] | ((unsigned long long)head[9] << 32);
      nvis = head[0];
      if (T != kTile && head[7]) {
        // maxCoverage truncates a splat at 16-px tiles: only they give the
        // reference's image
        last_cov16 = true;
        rerender_at16(sc, cam, bg);
        return;
      }
    }
    // a frame above the pair budget is cut into passes that each fit (below)
    if (pairs_all > budget() && train_mode)
      fail(DL_ERR_CAPACITY, "the differentiable render needs the frame's (tile, splat) pairs in one list (< 2^32)");
    if (keep_taps) {
      auto sid = to_host(sorted_ids, nvis, s);
      tap_append(taps[DL_TAP_SORTED_INDICES], sid.data(), nvis);
      auto sk = to_host(sorted_keys, nvis, s);
      tap_append(taps[DL_TAP_DEPTH_KEYS], sk.data(), nvis);
      auto rh = to_host(rec.as<float>(), (size_t)n * 12, s);
      auto ch = to_host(counts.as<uint32_t>(), n, s);
      std::vector<float> pr((size_t)nvis * 12);
      std::vector<uint32_t> pc(nvis);
      for (uint32_t j = 0; j < nvis; j++) {
        std::memcpy(&pr[(size_t)j * 12], &rh[(size_t)sid[j] * 12], 48);
        pc[j] = ch[sid[j]];
      }
      tap_append(taps[DL_TAP_PROJECTED], pr.data(), pr.size());
      tap_append(taps[DL_TAP_TILE_COUNTS], pc.data(), pc.size());
      tap_rank.assign(n, 0xFFFFFFFFu);
      for (uint32_t j = 0; j < nvis; j++) tap_rank[sid[j]] = j;
    }
  }
  stats.num_visible = nvis;

  const bool need_accum = keep_taps || train_mode;
  bool accum_valid = false;
  if (nvis == 0) {
    // pipeline.ts:210-212: background only
    rp.has_pairs = 0;
    rp.resume = 0;
    rp.finalize = 1;
    rp.write_accum = need_accum;
    if (need_accum) accum.reserve(npx * 16);
    mark(DL_STAGE_RASTERIZE);
    launch_rasterize(counting, T, G, units, s, nullptr, nullptr, nullptr, accum.as<float4>(), rgba.as<uint32_t>(), rp,
                     d_iters);
  } else {
    const bool span = opts.binning == 0 && span_binning_supported(tiles_x, tiles_y);
    // Front-to-back passes that stop filling closed tiles (span binning, no
    // taps). A pass costs about 30 launches, so by default they are used only
    // when the frame has enough pairs for the skipped binning to pay for them.
    const bool early_stop = span && !keep_taps && !train_mode && !opts.full_lists &&
                            (opts.first_pass || opts.pass_growth ||
                             (speculative ? last_pairs_all : pairs_all) >= kPassMinPairs);
    std::vector<uint32_t> cuts;  // pass starts, then nvis
    if (opts.chunk_cap && opts.chunk_cap < nvis && !train_mode) {
      for (uint64_t c0 = 0; c0 < nvis; c0 += opts.chunk_cap) cuts.push_back((uint32_t)c0);
    } else if (early_stop) {
      uint64_t len = opts.first_pass ? opts.first_pass : std::max<uint64_t>(kPassMinFirst, nvis / kPassFirstDiv);
      const uint64_t growth = opts.pass_growth ? opts.pass_growth : kPassGrowth;
      for (uint64_t c0 = 0; c0 < nvis; c0 += len, len *= growth) cuts.push_back((uint32_t)c0);
    } else {
      cuts.push_back(0);
    }
    cuts.push_back(nvis);
    // Over the pair budget (exact frames only: a speculative frame that
    // outgrows its capacity renders again exactly): every pass whose splats
    // hold more than cut_pairs() pairs is cut further, greedily in depth
    // order, from the per-splat tile counts. A pass's list never exceeds
    // its splats' counts (closed tiles only remove entries), so each fits.
    std::vector<uint64_t> cut_bound;  // pair bound of each pass (over budget only)
    if (!speculative && pairs_all > budget() && cuts.size() - 1 <= kMaxCutSegs) {
      // on the device; the host reads only the cut list
      const uint64_t cap = cut_pairs();
      const uint32_t nseg = (uint32_t)cuts.size() - 1;
      // greedy passes: two in a row hold more than cap, so each early-stop
      // pass of S pairs yields at most 2 S / cap + 1
      const uint32_t maxout = (uint32_t)std::min<uint64_t>(nseg + 2 * (pairs_all / cap) + 8, (uint64_t)nvis + nseg);
      const uint32_t nb = div_up(nvis, kCutBlock);
      const size_t p_bytes = (((size_t)nvis + 1) * 8 + 255) & ~(size_t)255;
      const size_t b_bytes = ((size_t)nb * 8 + 255) & ~(size_t)255;
      chunk_counts.reserve((size_t)nvis * 4);
      cut_ws.reserve(p_bytes + b_bytes + ((size_t)maxout * 2 + 1) * 8);
      auto *P = cut_ws.as<unsigned long long>();
      auto *bsum = reinterpret_cast<unsigned long long *>(cut_ws.as<uint8_t>() + p_bytes);
      auto *cut_out = reinterpret_cast<unsigned long long *>(cut_ws.as<uint8_t>() + p_bytes + b_bytes);
      CutSegs segs{};
      segs.n = nseg;
      for (uint32_t c = 0; c <= nseg; c++) segs.at[c] = cuts[c];
      k_cut_gather<<<nb, 256, 0, s>>>(counts.as<uint32_t>(), sorted_ids, nvis, chunk_counts.as<uint32_t>(), bsum);
      DL_HIP(hipGetLastError());
      k_cut_scan_blocks<<<1, 1024, 0, s>>>(bsum, nb);
      DL_HIP(hipGetLastError());
      k_cut_prefix<<<nb, 256, 0, s>>>(chunk_counts.as<uint32_t>(), bsum, nvis, P);
      DL_HIP(hipGetLastError());
      k_cut_greedy<<<1, 64, 0, s>>>(P, segs, cap, cut_out, maxout);
      DL_HIP(hipGetLastError());
      std::vector<unsigned long long> hc((size_t)maxout * 2 + 1);
      DL_HIP(hipMemcpyAsync(hc.data(), cut_out, hc.size() * 8, hipMemcpyDeviceToHost, s));
      DL_HIP(hipStreamSynchronize(s));
      if (hc[0] > maxout) fail(DL_ERR_INTERNAL, "pass cuts above their bound");
      std::vector<uint32_t> fine{0};
      for (uint64_t i = 0; i < hc[0]; i++) {
        fine.push_back((uint32_t)hc[1 + 2 * i]);
        cut_bound.push_back(hc[2 + 2 * i]);
      }
      if (fine.back() != nvis) fail(DL_ERR_INTERNAL, "pass cuts do not end at the visible count");
      cuts.swap(fine);
    } else if (!speculative && pairs_all > budget()) {
      const uint64_t cap = cut_pairs();
      chunk_counts.reserve((size_t)nvis * 4);
      k_gather_counts<<<div_up(nvis, 256), 256, 0, s>>>(counts.as<uint32_t>(), sorted_ids, nvis,
                                                        chunk_counts.as<uint32_t>());
      DL_HIP(hipGetLastError());
      const std::vector<uint32_t> hc = to_host(chunk_counts.as<uint32_t>(), nvis, s);
      std::vector<uint32_t> fine{0};
      for (size_t c = 0; c + 1 < cuts.size(); c++) {
        uint64_t run = 0;
        for (uint32_t j = cuts[c]; j < cuts[c + 1]; j++) {
          if (run > 0 && run + hc[j] > cap) {
            fine.push_back(j);
            cut_bound.push_back(run);
            run = 0;
          }
          run += hc[j];
        }
        fine.push_back(cuts[c + 1]);
        cut_bound.push_back(run);
      }
      cuts.swap(fine);
    }
    const uint32_t chunks = (uint32_t)cuts.size() - 1;
    // exact per-chunk pair totals (host sync) only where a pass needs them
    const bool exact_counts = !span || (chunks > 1 && !early_stop);
    // closed tiles are tracked across passes (they shrink later lists) also
    // when the passes come from the pair budget alone
    const bool track_closed = (early_stop || (!cut_bound.empty() && span && !keep_taps && !train_mode)) && chunks > 1;
    ranges.reserve(((size_t)num_tiles + 1) * 4);
    if (chunks > 1 || need_accum) accum.reserve(npx * 16);
    if (track_closed) {
      closed.reserve(flag_tiles);
      open_pre.reserve((size_t)tiles_y * (tiles_x + 1) * 2);
      // no clear: the first pass is not the last, reads no flag and writes
      // every tile's
    }
    const uint32_t key_bits = std::max<uint32_t>(1, bit_width_u32(num_tiles - 1));
    stats.pair_key_bits = key_bits;
    if (!exact_counts) stats.num_pairs = pairs_all;
    for (uint32_t c = 0; c < chunks; c++) {
      const uint32_t c0 = cuts[c];
      const uint32_t cn = cuts[c + 1] - c0;
      const bool last = c + 1 == chunks;
      const uint32_t *cids = sorted_ids + c0;
      stats.num_chunks++;
      // ---- pair count of this chunk (and scanned offsets for emit) ----
      if (exact_counts) mark(DL_STAGE_SCAN);
      uint64_t npairs64 = cut_bound.empty() ? pairs_all : cut_bound[c];  // bound of this chunk
      if (exact_counts) {
        chunk_counts.reserve((size_t)cn * 4);
        offsets.reserve((size_t)cn * 4);
        scan_ws.reserve(scan_workspace_bytes(cn));
        k_gather_counts<<<div_up(cn, 256), 256, 0, s>>>(counts.as<uint32_t>(), cids, cn, chunk_counts.as<uint32_t>());
        DL_HIP(hipGetLastError());
        exclusive_scan_u32(s, chunk_counts.as<uint32_t>(), offsets.as<uint32_t>(), cn, d_total, scan_ws.get());
        uint32_t t32 = 0;
        DL_HIP(hipMemcpyAsync(&t32, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        DL_HIP(hipStreamSynchronize(s));
        npairs64 = t32;
        stats.num_pairs += t32;
      }
      if (npairs64 > 0xFFFFFFF0ull) fail(DL_ERR_CAPACITY, "more than 2^32 (tile, splat) pairs in one chunk");
      const uint32_t npairs = (uint32_t)npairs64;
      rp.resume = accum_valid;
      rp.finalize = last;
      rp.write_accum = !last || need_accum;
      rp.closed = track_closed ? closed.as<uint8_t>() : nullptr;
      rp.final_on_close = track_closed && !need_accum;
      if (npairs == 0) {
        // pipeline.ts:306: nothing to rasterize in this chunk
        if (keep_taps) {
          std::vector<uint32_t> z(num_tiles + 1, 0u);
          tap_append(taps[DL_TAP_TILE_OFFSETS], z.data(), z.size());
        }
        if (last) {
          rp.has_pairs = 0;
          mark(DL_STAGE_RASTERIZE);
          launch_rasterize(counting, T, G, units, s, nullptr, nullptr, nullptr, accum.as<float4>(),
                           rgba.as<uint32_t>(), rp, d_iters);
        }
        continue;
      }
      pvals.reserve((size_t)npairs * 4);
      const uint32_t *svals = pvals.as<uint32_t>();
      if (span) {
        // stable two-level bucketing straight into per-tile lists (binning.hip)
        const LiveMap live = (track_closed && c > 0) ? LiveMap{open_pre.as<uint16_t>(), tiles_x + 1} : LiveMap{};
        const BinGrid grid{tiles_x, tiles_y, u.max_cov, u.tg};
        // first pass: its level-1 records and colours in one kernel
        const bool fused = defer_color && c == 0;
        if (fused) {
          mark(DL_STAGE_PROJECT);
          bin_ws.rects.reserve((size_t)std::max<uint32_t>(cn, 1) * 16);
          launch_rects_colors(s, sc, cids, cn, grid, rec.as<float4>(), bin_ws.rects.as<uint4>());
        }
        mark(DL_STAGE_EMIT);
        span_bin(s, rec.as<float4>(), cids, cn, grid, npairs, live, want_rect8 ? rect8.as<uint2>() : nullptr, n,
                 bin_ws, ranges.as<uint32_t>(), pvals.as<uint32_t>(), d_bin, d_entries, fused);
        if (tally) {
          const bool compacted = live.pre && want_rect8;
          k_bin_tally<<<1, 64, 0, s>>>(d_bin, compacted ? bin_ws.live_small.as<uint32_t>() : nullptr, cn, d_tally);
          DL_HIP(hipGetLastError());
        }
        if (defer_color && !fused) {
          mark(DL_STAGE_PROJECT);
          launch_colors(s, sc, nullptr, bin_ws.rects.as<uint4>(), cn, bin_ws.live_small.as<uint32_t>(),
                        rec.as<float4>());
        }
      } else {
        pkeys.reserve((size_t)npairs * 4);
        ptkeys.reserve((size_t)npairs * 4);
        ptvals.reserve((size_t)npairs * 4);
        psort_ws.reserve(radix_workspace_bytes(npairs, key_bits));
        mark(DL_STAGE_EMIT);
        k_emit<<<div_up(cn, 128), 128, 0, s>>>(rec.as<float4>(), cids, offsets.as<uint32_t>(), cn, u.tg, u.max_cov,
                                               pkeys.as<uint32_t>(), pvals.as<uint32_t>());
        DL_HIP(hipGetLastError());
        mark(DL_STAGE_PAIR_SORT);
        const bool in_tmp = radix_sort_pairs(s, pkeys.as<uint32_t>(), pvals.as<uint32_t>(), ptkeys.as<uint32_t>(),
                                             ptvals.as<uint32_t>(), npairs, (int)key_bits, psort_ws.get(),
                                             digit_bits());
        const uint32_t *skeys = in_tmp ? ptkeys.as<uint32_t>() : pkeys.as<uint32_t>();
        svals = in_tmp ? ptvals.as<uint32_t>() : pvals.as<uint32_t>();
        mark(DL_STAGE_RANGES);
        bin_ranges(s, skeys, npairs, num_tiles, ranges.as<uint32_t>());
        k_add_const<<<1, 64, 0, s>>>(npairs, d_entries);
        DL_HIP(hipGetLastError());
      }
      if (keep_taps) {
        // reference view: sorted tile ids and chunk-local depth ranks
        auto off = to_host(ranges.as<uint32_t>(), (size_t)num_tiles + 1, s);
        auto vals = to_host(svals, npairs, s);
        std::vector<uint32_t> keys_h(npairs);
        for (uint32_t t = 0; t < num_tiles; t++)
          for (uint32_t k = off[t]; k < off[t + 1]; k++) keys_h[k] = t;
        for (auto &v : vals) v = tap_rank[v] - c0;
        tap_append(taps[DL_TAP_PAIR_TILES], keys_h.data(), keys_h.size());
        tap_append(taps[DL_TAP_PAIR_SPLATS], vals.data(), vals.size());
        tap_append(taps[DL_TAP_TILE_OFFSETS], off.data(), off.size());
      }
      rp.has_pairs = 1;
      train_vals = svals;
      mark(DL_STAGE_RASTERIZE);
      RastParams rpc = rp;
      const bool ordered = c == 0 && units >= kOrderMinTiles;
      if (ordered) {
        if (next_order && tile_key_tiles == units && tile_key_shape == shape) {
          // the order sorted from the last frame's keys; the wait also keeps
          // this launch from overwriting the keys before that sort read them
          DL_HIP(hipStreamWaitEvent(s, order_done, 0));
          rpc.order = next_order;
        }
        tile_key.reserve((size_t)units * 4);
        rpc.cost_key = tile_key.as<uint32_t>();
      }
      launch_rasterize(counting, T, G, units, s, rec.as<float4>(), ranges.as<uint32_t>(), svals, accum.as<float4>(),
                       rgba.as<uint32_t>(), rpc, d_iters);
      if (ordered) {
        // sort this frame's keys into the next frame's order on the second
        // stream: one stable 6-bit pass whose values are the tile indices
        if (!order_stream) {
          DL_HIP(hipStreamCreateWithFlags(&order_stream, hipStreamNonBlocking));
          DL_HIP(hipEventCreateWithFlags(&keys_done, hipEventDisableTiming));
          DL_HIP(hipEventCreateWithFlags(&order_done, hipEventDisableTiming));
        }
        tile_key_tmp.reserve((size_t)units * 4);
        tile_order.reserve((size_t)units * 4);
        tile_order_tmp.reserve((size_t)units * 4);
        tile_sort_ws.reserve(radix_workspace_bytes(units, 6));
        DL_HIP(hipEventRecord(keys_done, s));
        DL_HIP(hipStreamWaitEvent(order_stream, keys_done, 0));
        const bool in_tmp = radix_sort_pairs(order_stream, tile_key.as<uint32_t>(), tile_order.as<uint32_t>(),
                                             tile_key_tmp.as<uint32_t>(), tile_order_tmp.as<uint32_t>(), units, 6,
                                             tile_sort_ws.get(), 6, RadixKeyMap{}, true);
        DL_HIP(hipEventRecord(order_done, order_stream));
        next_order = in_tmp ? tile_order_tmp.as<uint32_t>() : tile_order.as<uint32_t>();
        tile_key_tiles = units;
        tile_key_shape = shape;
      }
      accum_valid = true;
      if (track_closed && !last) {
        if (T == 32)
          k_open_prefix_quads<<<tiles_y, 256, 0, s>>>(closed.as<uint8_t>(), tg.t16x, tg.t16y, tiles_x,
                                                      open_pre.as<uint16_t>());
        else
          k_open_prefix<<<tiles_y, 256, 0, s>>>(closed.as<uint8_t>(), tiles_x, open_pre.as<uint16_t>());
        DL_HIP(hipGetLastError());
      }
    }
  }
  mark(-1);
  {
    // small: [0] visible, [2..3] key span, [7] coverage flag, [8..9] all
    // pairs, [10..15] loop counters, [16..21] tallies
    uint32_t sm[22] = {};
    static_assert(sizeof(sm) <= kHostSmallBytes, "pinned readback too small");
    void *hs = pinned_small();
    DL_HIP(hipMemcpyAsync(hs, d_nvis, sizeof(sm), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    std::memcpy(sm, hs, sizeof(sm));
    auto u64_at = [&](int k) { return (unsigned long long)sm[k] | ((unsigned long long)sm[k + 1] << 32); };
    if (counting) {
      stats.pixel_iters = u64_at(10);
      stats.pairs_consumed = u64_at(12);
    }
    if (tally) {
      stats.binned_splats = u64_at(16);
      stats.row_spans = u64_at(18);
      stats.group_spans = u64_at(20);
    }
    stats.list_entries = u64_at(14);
    if (n > 0) {
      const unsigned long long all = u64_at(8);
      last_pairs_all = all;
      // frames over the pair budget stay exact (their passes are cut from
      // the per-splat counts, which a speculative frame does not write)
      if (all <= budget())
        spec_cap = std::max<uint64_t>(spec_cap, std::min<uint64_t>(all + all / 8 + 4096, budget()));
      else
        spec_cap = 0;
      // span of the visible depth keys: (kmax - kmin + 1) and the sentinel
      // must fit the sort's bits
      bool keys_fit = true;
      if (sm[0] > 0 && sm[2] <= sm[3]) {
        const uint64_t span = (uint64_t)sm[3] - (uint64_t)sm[2] + 1ull;
        uint32_t need = 0;
        while (need < 33 && (span >> need) != 0) need++;
        keys_fit = key_bits_used == 0 || need <= key_bits_used;
        if (!keys_fit) key_headroom = std::min<uint32_t>(2, key_headroom + 1);
        spec_key_bits = std::min<uint32_t>(32, need + key_headroom);
      }
      stats.coverage_truncated = sm[7] ? 1u : 0u;
      if (T != kTile && sm[7]) {
        last_cov16 = true;
        rerender_at16(sc, cam, bg);
        return;
      }
      // an exact re-render below keeps this frame's tile size; the next
      // frame starts from what this one found
      struct SetCov {
        bool &f;
        bool v;
        ~SetCov() { f = v; }
      } set_cov{last_cov16, sm[7] != 0};
      if (speculative) {
        stats.num_visible = sm[0];
        stats.num_pairs = all;
        if (all > pairs_all || !keys_fit) {
          // the lists were cut at the capacity, or the depth keys spanned
          // more bits than the sort used: render again with exact totals
          struct Reset {
            bool &f;
            ~Reset() { f = false; }
          } reset{force_exact};
          force_exact = true;
          render(sc, cam, bg);
          return;
        }
      }
    }
  }
  DL_HIP(hipStreamSynchronize(s));
  if (keep_taps) {
    auto acc = to_host(accum.as<float>(), npx * 4, s);
    tap_append(taps[DL_TAP_ACCUM], acc.data(), acc.size());
  }
  if (timing && ev_used > 1) {
    for (size_t k = 0; k + 1 < ev_used; k++) {
      float ms = 0.0f;
      DL_HIP(hipEventElapsedTime(&ms, events[k], events[k + 1]));
      if (ev_stage[k] >= 0 && ev_stage[k] < DL_NUM_STAGES) stats.stage_ms[ev_stage[k]] += ms;
      stats.total_ms += ms;
    }
  }
}

void Renderer::rerender_at16(const SceneView &sc, const dl_camera &cam, const dl_background &bg) {
  struct Reset {
    bool &f;
    ~Reset() { f = false; }
  } reset{force_tile16};
  force_tile16 = true;
  render(sc, cam, bg);
  stats.coverage_truncated = 1;
}

void Renderer::read_image(uint8_t *out) {
  if (!width) fail(DL_ERR_INVALID, "no image has been rendered yet");
  ctx->make_current();
  ctx->staging.to_host(out, rgba.get(), (size_t)width * height * 4, ctx->stream);
}

Renderer::~Renderer() {
  for (auto e : events) (void)hipEventDestroy(e);
  if (host_small) (void)hipHostFree(host_small);
  if (order_stream) {
    (void)hipStreamSynchronize(order_stream);
    (void)hipStreamDestroy(order_stream);
    (void)hipEventDestroy(keys_done);
    (void)hipEventDestroy(order_done);
  }
}

}  // namespace dl
