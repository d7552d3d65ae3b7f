// Batched views on gfx950: many (scene, camera) views through one pass of
// every stage, captured once as a hipGraph and replayed.
//
// A group of views is laid out view-major: the items of view v are the
// splats of its scene, at item0 .. item0 + n - 1 of the group's arrays. Each
// stage is one launch over the whole group:
//   k_batch_cull    f64 cull, depth key and projection of every item
//                   (cull_project_one, the single-view code), 256 items of
//                   one view per block; per-block (tile, splat) pair counts
//   k_sum_pairs     the group's pair total, checked against the capacity
//   depth sort      stable segmented LSD radix sort, one segment per view
//   binning         span binning (binning.hip) over a virtual image that
//                   stacks the tile rows of all views of the group
//   k_batch_raster  one wave per tile of every view (rasterize_tile)
// Every view's image is the single-view renderer's image of that (scene,
// camera), bit for bit: the same f32 arithmetic on the same per-tile lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "batch.h"
#include "dl_common.h"
#include "render_device.h"

namespace dl {

namespace {

constexpr uint32_t kItemBlock = 256;
static_assert(kItemBlock == (uint32_t)kB1, "level-1 blocks are counted in items of the level-1 kernels' block");

// view owning block b (views ordered by first block); views with no blocks
// never own one because the search takes the last view with blk0 <= b
__device__ inline uint32_t view_of_block(const BView *__restrict__ views, uint32_t nv, uint32_t b, bool tiles) {
  uint32_t lo = 0, hi = nv - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    const uint32_t first = tiles ? views[mid].tile_blk0 : views[mid].blk0;
    if (first <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kItemBlock) void k_batch_cull(const BView *__restrict__ views, uint32_t nv,
                                                           const SceneView *__restrict__ scenes,
                                                           uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                           float4 *__restrict__ rec,
                                                           unsigned long long *__restrict__ parts,
                                                           uint2 *__restrict__ rect8, uint32_t max_cov) {
  __shared__ unsigned long long s_pairs[kItemBlock / 64];
  const BView &V = views[view_of_block(views, nv, blockIdx.x, false)];
  const uint32_t i = (blockIdx.x - V.blk0) * kItemBlock + threadIdx.x;
  uint32_t count = 0;
  if (i < V.n) {
    const SceneView sc = scenes[V.scene];
    const ProjUniforms u = V.u;
    const size_t o = (size_t)V.item0 + i;
    bool vis = false, cov16 = false;
    switch (sc.bands) {
      case 0: cull_project_one<0>(sc, u, V.near_plane, V.far_plane, i, o, (uint32_t)o, keys, vals, rec, nullptr, vis, count, cov16); break;
      case 1: cull_project_one<1>(sc, u, V.near_plane, V.far_plane, i, o, (uint32_t)o, keys, vals, rec, nullptr, vis, count, cov16); break;
      case 2: cull_project_one<2>(sc, u, V.near_plane, V.far_plane, i, o, (uint32_t)o, keys, vals, rec, nullptr, vis, count, cov16); break;
      default: cull_project_one<3>(sc, u, V.near_plane, V.far_plane, i, o, (uint32_t)o, keys, vals, rec, nullptr, vis, count, cov16); break;
    }
    // the packed rect in the view's grid (rows from the view's first), for
    // the live tests of later passes
    if (rect8) rect8[o] = pack_rect8(splat_rect(rec[3 * o], grid16(V.tiles_x, V.tiles_y), max_cov));
  }
  unsigned long long pc = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) pc += __shfl_xor(pc, d, 64);
  if (__lane_id() == 0) s_pairs[threadIdx.x >> 6] = pc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < kItemBlock / 64; w++) t += s_pairs[w];
    parts[blockIdx.x] = t;
  }
}

// ---- live-splat compaction of a pass after the first, per view (the
// single-view renderer's k_live_bits / k_live_count / k_rects_live) ----

// bit i of view V's words (from V.bit0): item V.item0 + i touches an open
// tile of the stacked image (rect_live)
__global__ __launch_bounds__(kItemBlock) void k_batch_live_bits(const BView *__restrict__ views, uint32_t nv,
                                                                const uint2 *__restrict__ rect8, LiveMap live,
                                                                unsigned long long *__restrict__ bits) {
  const BView &V = views[view_of_block(views, nv, blockIdx.x, false)];
  const uint32_t i = (blockIdx.x - V.blk0) * kItemBlock + threadIdx.x;
  bool lv = false;
  if (i < V.n) {
    Rect R = unpack_rect8(rect8[(size_t)V.item0 + i]);
    R.y0 += V.row0;
    lv = rect_live(R, live);
  }
  const unsigned long long b = __ballot(lv);
  const uint32_t first = i - (uint32_t)__lane_id();
  if (__lane_id() == 0 && first < V.n) bits[(size_t)V.bit0 + (first >> 6)] = b;
}

// index of the level-1 view that owns block b (views ordered by first block)
__device__ inline uint32_t l1_index(const L1View *__restrict__ tab, uint32_t ntab, uint32_t b) {
  uint32_t lo = 0, hi = ntab - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (tab[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline bool item_live(const unsigned long long *__restrict__ bits, const L1View &v, uint32_t o) {
  const uint32_t li = o - v.vitem0;
  return (bits[(size_t)v.bit0 + (li >> 6)] >> (li & 63u)) & 1ull;
}

// live items of every level-1 block of the pass
__global__ __launch_bounds__(kItemBlock) void k_batch_live_count(const uint32_t *__restrict__ sorted,
                                                                 const L1View *__restrict__ tab, uint32_t ntab,
                                                                 const unsigned long long *__restrict__ bits,
                                                                 uint32_t *__restrict__ cnt) {
  __shared__ uint32_t s_w[kItemBlock / 64];
  const L1View v = tab[l1_index(tab, ntab, blockIdx.x)];
  const uint32_t jj = (blockIdx.x - v.blk0) * kItemBlock + threadIdx.x;
  const bool lv = jj < v.n && item_live(bits, v, sorted[(size_t)v.item0 + jj]);
  const unsigned long long b = __ballot(lv);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// The live items of each view moved, in depth order, to the front of the
// view's part of the pass as level 1 reads them (id and rect), and the
// view's live count into its table entry (n_live); pre is the exclusive
// scan of the block counts with the total after it.
__global__ __launch_bounds__(kItemBlock) void k_batch_rects_live(const uint32_t *__restrict__ sorted,
                                                                 L1View *__restrict__ tab, uint32_t ntab,
                                                                 const unsigned long long *__restrict__ bits,
                                                                 const uint2 *__restrict__ rect8,
                                                                 const uint32_t *__restrict__ pre,
                                                                 uint4 *__restrict__ rects) {
  __shared__ uint32_t s_w[kItemBlock / 64];
  const uint32_t vi = l1_index(tab, ntab, blockIdx.x);
  const uint32_t item0 = tab[vi].item0, n = tab[vi].n, blk0 = tab[vi].blk0;
  L1View v{};
  v.vitem0 = tab[vi].vitem0;
  v.bit0 = tab[vi].bit0;
  const uint32_t kb = blockIdx.x - blk0;
  const uint32_t jj = kb * kItemBlock + threadIdx.x;
  uint32_t o = 0;
  bool lv = false;
  if (jj < n) {
    o = sorted[(size_t)item0 + jj];
    lv = item_live(bits, v, o);
  }
  const unsigned long long b = __ballot(lv);
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) s_w[wave] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t pos = pre[blockIdx.x] - pre[blk0];
  for (int w = 0; w < wave; w++) pos += s_w[w];
  pos += (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  if (lv) rects[(size_t)item0 + pos] = pack_rect(o, unpack_rect8(rect8[o]));
  if (kb == 0 && threadIdx.x == 0) tab[vi].n_live = pre[blk0 + (n + kItemBlock - 1) / kItemBlock] - pre[blk0];
}

__global__ __launch_bounds__(1024) void k_sum_pairs(const unsigned long long *__restrict__ parts, uint32_t n,
                                                    unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_w[16];
  unsigned long long p = 0;
  for (uint32_t k = threadIdx.x; k < n; k += 1024) p += parts[k];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) p += __shfl_xor(p, d, 64);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 16; w++) t += s_w[w];
    *out = t;
  }
}

// One pass over every tile of the group. resume / finalize / write_accum and
// the closed flags follow the single-view renderer's passes; accum and the
// images are indexed by the view's first pixel.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_batch_raster(const BView *__restrict__ views, uint32_t nv, uint32_t cols,
                                                     const float4 *__restrict__ rec,
                                                     const uint32_t *__restrict__ ranges,
                                                     const uint32_t *__restrict__ pvals, float4 *__restrict__ accum,
                                                     uint32_t *__restrict__ images, uint8_t *__restrict__ closed,
                                                     uint32_t resume, uint32_t finalize) {
  __shared__ RastLds lds;
  const BView &V = views[view_of_block(views, nv, blockIdx.x, true)];
  const uint32_t t = blockIdx.x - V.tile_blk0;
  const uint32_t tx = t % V.tiles_x, ty = t / V.tiles_x;
  const uint32_t vtile = (V.row0 + ty) * cols + tx;
  RastParams p{};
  p.tiles_x = V.tiles_x;
  p.width = V.width;
  p.height = V.height;
  p.has_pairs = 1;
  p.resume = resume;
  p.finalize = finalize;
  p.write_accum = !finalize;
  p.closed = closed;
  p.final_on_close = closed != nullptr;
  p.bg_r = V.bg[0];
  p.bg_g = V.bg[1];
  p.bg_b = V.bg[2];
  p.bg_a = V.bg[3];
  rasterize_tile<false>(tx, ty, vtile, ranges[vtile], ranges[vtile + 1], rec, pvals, accum ? accum + V.px0 : nullptr,
                        images + V.px0, p, nullptr, lds);
}

template <class T>
void upload(DeviceBuffer &dst, const std::vector<T> &src, hipStream_t s) {
  dst.reserve(std::max<size_t>(src.size(), 1) * sizeof(T));
  if (!src.empty()) DL_HIP(hipMemcpyAsync(dst.get(), src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace

Batch::~Batch() {
  drop_graph();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
}

void Batch::drop_graph() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

void Batch::plan(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews) {
  std::vector<uint64_t> lay;
  lay.push_back(scenes.size());
  for (const auto &sc : scenes) lay.push_back(((uint64_t)sc.count << 8) | sc.bands);
  lay.push_back(nviews);
  for (uint32_t v = 0; v < nviews; v++) {
    const dl_camera &c = views[v].camera;
    if (views[v].scene >= scenes.size()) fail(DL_ERR_INVALID, "view " + std::to_string(v) + ": scene index out of range");
    if (c.width == 0 || c.height == 0) fail(DL_ERR_INVALID, "view " + std::to_string(v) + ": width and height must be > 0");
    if (c.projection != DL_PINHOLE && c.projection != DL_EQUIRECT)
      fail(DL_ERR_INVALID, "view " + std::to_string(v) + ": unknown projection");
    if (!span_binning_supported(div_up(c.width, kTile), div_up(c.height, kTile)))
      fail(DL_ERR_INVALID, "view " + std::to_string(v) + ": batched views are limited to 4096 x 16384 pixels");
    lay.push_back(((uint64_t)views[v].scene << 40) | ((uint64_t)c.width << 20) | c.height);
  }
  for (const auto &sc : scenes)
    if (sc.bands > 3) fail(DL_ERR_INVALID, "shBands must be 0..3");
  if (lay != break_layout) {
    // group splits and solo views were found for another layout
    break_layout = lay;
    view_break.assign(nviews, 0);
    view_solo.assign(nviews, 0);
  }
  if (lay == layout && !groups.empty()) return;

  // a new layout: new groups, tables and buffers; the old graph is invalid
  drop_graph();
  layout = lay;
  groups.clear();
  hviews.assign(nviews, BView{});
  hl1.clear();
  hseg.clear();
  px_total = 0;
  items_max = 0;
  nv_max = 0;
  passes_max = 1;
  uint32_t blocks_max = 0, pass_blocks_max = 0, words_max = 0;
  size_t rows_max = 0, sat_max = 0, csat_max = 0;
  vtiles_max = 0;
  // front-to-back passes of a view of n splats, on the single-view
  // renderer's schedule (one pass with full_lists); the batch always uses
  // them because its launches serve every view of the group at once
  const uint64_t growth = opts.pass_growth ? opts.pass_growth : kPassGrowth;
  auto cuts_of = [&](uint32_t n) {
    std::vector<uint32_t> c{0};
    if (!opts.full_lists) {
      uint64_t len = opts.first_pass ? opts.first_pass : std::max<uint64_t>(kPassMinFirst, n / kPassFirstDiv);
      for (uint64_t c0 = len; c0 < n; len *= growth, c0 += len) c.push_back((uint32_t)c0);
    }
    c.push_back(n);
    return c;
  };
  for (uint32_t v = 0; v < nviews;) {
    if (view_solo[v]) {
      // rendered by the single-view renderer: only its image slot here
      BView &b = hviews[v];
      b.scene = views[v].scene;
      b.n = scenes[views[v].scene].count;
      b.width = views[v].camera.width;
      b.height = views[v].camera.height;
      b.tiles_x = div_up(b.width, kTile);
      b.tiles_y = div_up(b.height, kTile);
      b.px0 = px_total;
      px_total += (uint64_t)b.width * b.height;
      v++;
      continue;
    }
    BatchGroup g;
    g.v0 = v;
    g.seg_off = hseg.size();
    uint64_t items = 0;
    std::vector<std::vector<uint32_t>> cuts;
    while (v < nviews && !view_solo[v] && (g.nv == 0 || (!view_break[v] && items + scenes[views[v].scene].count <= item_cap))) {
      const uint32_t n = scenes[views[v].scene].count;
      const dl_camera &c = views[v].camera;
      BView &b = hviews[v];
      b.scene = views[v].scene;
      b.n = n;
      b.item0 = (uint32_t)items;
      b.blk0 = g.cull_blocks;
      b.tile_blk0 = g.tiles;
      b.row0 = g.rows;
      b.width = c.width;
      b.height = c.height;
      b.tiles_x = div_up(c.width, kTile);
      b.tiles_y = div_up(c.height, kTile);
      b.px0 = px_total;
      b.bit0 = g.bit_words;
      g.bit_words += div_up(n, 64);
      hseg.push_back(b.item0);
      cuts.push_back(cuts_of(n));
      items += n;
      g.cull_blocks += div_up(n, kItemBlock);
      g.tiles += b.tiles_x * b.tiles_y;
      g.rows += b.tiles_y;
      g.cols = std::max(g.cols, b.tiles_x);
      g.max_view_rows = std::max(g.max_view_rows, b.tiles_y);
      px_total += (uint64_t)c.width * c.height;
      g.nv++;
      v++;
    }
    if (items > 0xFFFFFFF0ull) fail(DL_ERR_CAPACITY, "a single view's scene exceeds 2^32 splats");
    g.items = (uint32_t)items;
    hseg.push_back(g.items);
    size_t passes = 1;
    for (const auto &c : cuts) passes = std::max(passes, c.size() - 1);
    for (size_t p = 0; p < passes; p++) {
      g.l1_off.push_back(hl1.size());
      uint32_t blocks = 0;
      size_t n1 = 0;
      for (uint32_t k = 0; k < g.nv; k++) {
        const BView &b = hviews[g.v0 + k];
        const auto &c = cuts[k];
        const uint32_t a0 = p + 1 < c.size() ? c[p] : b.n, a1 = p + 1 < c.size() ? c[p + 1] : b.n;
        L1View l;
        l.item0 = b.item0 + a0;
        l.n = a1 - a0;
        l.blk0 = blocks;
        l.row0 = b.row0;
        l.rows = b.tiles_y;
        l.c1_base = (uint32_t)n1;
        l.tiles_x = b.tiles_x;
        l.max_cov = max_cov;
        l.n_live = l.n;
        l.vitem0 = b.item0;
        l.bit0 = b.bit0;
        hl1.push_back(l);
        const uint32_t nb = div_up(l.n, kItemBlock);
        blocks += nb;
        n1 += (size_t)b.tiles_y * nb;
      }
      if (n1 > 0xFFFFFFF0ull) fail(DL_ERR_CAPACITY, "batch group too large");
      g.blocks.push_back(blocks);
      g.n1.push_back(n1);
      pass_blocks_max = std::max(pass_blocks_max, blocks);
    }
    items_max = std::max(items_max, g.items);
    nv_max = std::max(nv_max, g.nv);
    passes_max = std::max<uint32_t>(passes_max, (uint32_t)passes);
    blocks_max = std::max(blocks_max, g.cull_blocks);
    words_max = std::max(words_max, g.bit_words);
    vtiles_max = std::max(vtiles_max, (size_t)g.rows * g.cols);
    rows_max = std::max<size_t>(rows_max, g.rows);
    sat_max = std::max<size_t>(sat_max, ((size_t)g.rows + 1) * (g.cols + 1));
    csat_max = std::max<size_t>(csat_max, ((size_t)(g.rows + 3) / 4 + 1) * ((g.cols + 3) / 4 + 1));
    groups.push_back(g);
  }
  hpairs.assign(groups.size(), 0);
  hscenes = scenes;

  hipStream_t s = ctx->stream;
  upload(d_l1, hl1, s);
  upload(d_seg, hseg, s);
  d_views.reserve(std::max<size_t>(nviews, 1) * sizeof(BView));
  d_scenes.reserve(std::max<size_t>(scenes.size(), 1) * sizeof(SceneView));
  d_parts.reserve(std::max<size_t>(blocks_max, 1) * 8);
  d_pairs.reserve(groups.size() * 8 + 8);
  const size_t im = std::max<size_t>(items_max, 1);
  keys.reserve(im * 4);
  vals.reserve(im * 4);
  tkeys.reserve(im * 4);
  tvals.reserve(im * 4);
  rec.reserve(im * 48);
  rects.reserve(im * 16);
  sort_ws.reserve(seg_radix_workspace_bytes(im, std::max<uint32_t>(nv_max, 1)));
  ranges.reserve((vtiles_max + 1) * 4);
  images.reserve(std::max<uint64_t>(px_total, 1) * 4);
  small.reserve(256);
  if (passes_max > 1) {
    accum.reserve(std::max<uint64_t>(px_total, 1) * 16);
    closed.reserve(std::max<size_t>(vtiles_max, 1));
    open_pre.reserve(std::max<size_t>(vtiles_max + rows_max, 1) * 2);
    live_sat.reserve(sat_max * 4);
    live_csat.reserve(csat_max * 4);
    rect8.reserve(im * 8);
    live_bits.reserve(((size_t)words_max + 1) * 8);
    live_cnt.reserve(((size_t)pass_blocks_max + 1) * 4);
    live_pre.reserve(((size_t)pass_blocks_max + 1) * 4);
    live_scan_ws.reserve(scan_workspace_bytes(std::max<uint32_t>(pass_blocks_max, 1)));
  }
  DL_HIP(hipStreamSynchronize(s));
  stats.num_views = nviews;
  stats.num_groups = (uint32_t)groups.size();
}

// All launches of one call, in stream order. eager: the group's pair total
// is read back after the cull (one sync per group) and the list capacity
// grows to fit it; otherwise every launch is sized by the current plan, so
// the sequence can be captured (overflow then clamps and is caught after
// the call).
void Batch::enqueue(bool eager) {
  hipStream_t s = ctx->stream;
  const SceneView *d_sc = d_scenes.as<SceneView>();
  unsigned long long *d_gp = d_pairs.as<unsigned long long>();
  for (size_t gi = 0; gi < groups.size(); gi++) {
    BatchGroup &g = groups[gi];
    const BView *gv = d_views.as<BView>() + g.v0;
    if (g.cull_blocks) {
      // the values are the item indices: the sort's first pass writes them
      k_batch_cull<<<g.cull_blocks, kItemBlock, 0, s>>>(gv, g.nv, d_sc, keys.as<uint32_t>(), nullptr,
                                                        rec.as<float4>(), d_parts.as<unsigned long long>(),
                                                        g.blocks.size() > 1 ? rect8.as<uint2>() : nullptr, max_cov);
      DL_HIP(hipGetLastError());
    }
    k_sum_pairs<<<1, 1024, 0, s>>>(d_parts.as<unsigned long long>(), g.cull_blocks, d_gp + gi);
    DL_HIP(hipGetLastError());
    if (eager) {
      unsigned long long need = 0;
      DL_HIP(hipMemcpyAsync(&need, d_gp + gi, sizeof(need), hipMemcpyDeviceToHost, s));
      DL_HIP(hipStreamSynchronize(s));
      if (need > group_pairs()) throw GroupOverflow{gi};
      if (need > g.pair_cap || g.pair_cap == 0)
        g.pair_cap = (uint32_t)std::min<uint64_t>(group_pairs(), need + need / 8 + 4096);
      pvals.reserve((size_t)g.pair_cap * 4);
    }
    const bool in_tmp = seg_radix_sort_pairs(s, keys.as<uint32_t>(), vals.as<uint32_t>(), tkeys.as<uint32_t>(),
                                             tvals.as<uint32_t>(), d_seg.as<uint32_t>() + g.seg_off, g.nv, g.items,
                                             32, opts.sort_digit_bits ? (int)std::min<uint32_t>(opts.sort_digit_bits, 8) : 8,
                                             sort_ws.get(), true);
    const uint32_t *sorted = in_tmp ? tvals.as<uint32_t>() : vals.as<uint32_t>();
    const uint32_t passes = (uint32_t)g.blocks.size();
    const size_t vtiles = (size_t)g.rows * g.cols;
    if (passes > 1) {
      k_fill_u8<<<div_up(vtiles, 256), 256, 0, s>>>(closed.as<uint8_t>(), vtiles, 0);
      DL_HIP(hipGetLastError());
    }
    for (uint32_t p = 0; p < passes; p++) {
      L1View *tab = d_l1.as<L1View>() + g.l1_off[p];
      if (p == 0) {
        batch_rects(s, rec.as<float4>(), sorted, tab, g.nv, g.blocks[p], rects.as<uint4>());
      } else if (g.blocks[p]) {
        // only the splats that touch an open tile, compacted per view
        const LiveMap lm = open_sat_tall(s, open_pre.as<uint16_t>(), g.rows, g.cols, live_sat.as<uint32_t>(),
                                         live_csat.as<uint32_t>());
        unsigned long long *bits = live_bits.as<unsigned long long>();
        k_batch_live_bits<<<g.cull_blocks, kItemBlock, 0, s>>>(gv, g.nv, rect8.as<uint2>(), lm, bits);
        DL_HIP(hipGetLastError());
        k_batch_live_count<<<g.blocks[p], kItemBlock, 0, s>>>(sorted, tab, g.nv, bits, live_cnt.as<uint32_t>());
        DL_HIP(hipGetLastError());
        exclusive_scan_u32(s, live_cnt.as<uint32_t>(), live_pre.as<uint32_t>(), g.blocks[p],
                           live_pre.as<uint32_t>() + g.blocks[p], live_scan_ws.get());
        k_batch_rects_live<<<g.blocks[p], kItemBlock, 0, s>>>(sorted, tab, g.nv, bits, rect8.as<uint2>(),
                                                              live_pre.as<uint32_t>(), rects.as<uint4>());
        DL_HIP(hipGetLastError());
      }
      SpanBinArgs a;
      a.rects = rects.as<uint4>();
      a.tab = tab;
      a.ntab = g.nv;
      a.l1_blocks = g.blocks[p];
      a.n1 = g.n1[p];
      a.rows = g.rows;
      a.cols = g.cols;
      a.max_view_rows = g.max_view_rows;
      a.pair_cap = g.pair_cap;
      if (p > 0) a.live = LiveMap{open_pre.as<uint16_t>(), g.cols + 1};
      span_bin_rects(s, a, bin_ws, ranges.as<uint32_t>(), pvals.as<uint32_t>(), small.as<uint32_t>());
      const bool last = p + 1 == passes;
      if (g.tiles) {
        k_batch_raster<<<g.tiles, 64, 0, s>>>(gv, g.nv, g.cols, rec.as<float4>(), ranges.as<uint32_t>(),
                                              pvals.as<uint32_t>(), passes > 1 ? accum.as<float4>() : nullptr,
                                              images.as<uint32_t>(), passes > 1 ? closed.as<uint8_t>() : nullptr,
                                              p > 0 ? 1u : 0u, last ? 1u : 0u);
        DL_HIP(hipGetLastError());
      }
      if (!last) {
        k_open_prefix<<<g.rows, 256, 0, s>>>(closed.as<uint8_t>(), g.cols, open_pre.as<uint16_t>());
        DL_HIP(hipGetLastError());
      }
    }
  }
}

void Batch::render(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews) {
  if (nviews == 0) fail(DL_ERR_INVALID, "no views");
  ctx->make_current();
  hipStream_t s = ctx->stream;
  if (!ev0) DL_HIP(hipEventCreate(&ev0));
  if (!ev1) DL_HIP(hipEventCreate(&ev1));
  for (uint32_t attempt = 0;; attempt++) {
    plan(scenes, views, nviews);
    try {
      run_groups(scenes, views, nviews);
      break;
    } catch (const GroupOverflow &o) {
      // split the group at its middle view, or render its one view solo,
      // and plan again (the launches already queued finish first: the new
      // plan reallocates their buffers)
      DL_HIP(hipStreamSynchronize(s));
      const BatchGroup &g = groups[o.group];
      if (g.nv > 1)
        view_break[g.v0 + g.nv / 2] = 1;
      else
        view_solo[g.v0] = 1;
      layout.clear();
      drop_graph();
      if (attempt > 2 * nviews) fail(DL_ERR_INTERNAL, "batch groups could not be split to fit their pairs");
    }
  }
  // solo views: the single-view renderer, into their image slots
  for (uint32_t v = 0; v < nviews; v++) {
    if (!view_solo[v]) continue;
    if (!solo_r) {
      solo_r.reset(new Renderer());
      solo_r->ctx = ctx;
    }
    solo_r->opts = opts;
    solo_r->opts.chunk_cap = 0;
    solo_r->max_cov = max_cov;
    solo_r->pair_budget = pair_budget;
    solo_r->render(scenes[views[v].scene], views[v].camera, views[v].background);
    const BView &b = hviews[v];
    DL_HIP(hipMemcpyAsync(images.as<uint32_t>() + b.px0, solo_r->rgba.get(), (size_t)b.width * b.height * 4,
                          hipMemcpyDeviceToDevice, s));
    stats.num_pairs += solo_r->stats.num_pairs;
  }
  DL_HIP(hipStreamSynchronize(s));
}

// The groups' launches of one call: the cameras and backgrounds go into the
// view table, then the captured graph replays (or the launches run
// directly and are captured). Throws GroupOverflow from a direct run.
void Batch::run_groups(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews) {
  hipStream_t s = ctx->stream;
  // cameras and backgrounds of this call
  for (uint32_t v = 0; v < nviews; v++) {
    BView &b = hviews[v];
    const dl_camera &c = views[v].camera;
    b.u = make_uniforms(c, max_cov);
    b.near_plane = c.near_plane;
    b.far_plane = c.far_plane;
    b.bg[0] = views[v].background.r;
    b.bg[1] = views[v].background.g;
    b.bg[2] = views[v].background.b;
    b.bg[3] = views[v].background.a;
  }
  hscenes = scenes;
  upload(d_views, hviews, s);
  upload(d_scenes, hscenes, s);

  auto run = [&](bool replay) {
    DL_HIP(hipEventRecord(ev0, s));
    if (replay) {
      DL_HIP(hipGraphLaunch(exec, s));
    } else {
      enqueue(true);
    }
    DL_HIP(hipEventRecord(ev1, s));
    DL_HIP(hipMemcpyAsync(hpairs.data(), d_pairs.get(), hpairs.size() * 8, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  };
  const bool replay = use_graph && exec != nullptr;
  run(replay);
  if (replay) {
    stats.graph_replays++;
    bool overflow = false;
    for (size_t gi = 0; gi < groups.size(); gi++)
      if (hpairs[gi] > groups[gi].pair_cap) overflow = true;
    if (overflow) {
      // new cameras need longer lists than the captured plan holds
      drop_graph();
      run(false);
    }
  }
  float ms = 0.0f;
  DL_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  stats.device_ms = ms;
  stats.num_pairs = 0;
  for (auto p : hpairs) stats.num_pairs += p;
  if (use_graph && !exec && !groups.empty()) {
    DL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    DL_HIP(hipStreamEndCapture(s, &graph));
    DL_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    stats.captures++;
  }
}

void Batch::read_image(uint32_t view, uint8_t *out) const {
  if (view >= hviews.size()) fail(DL_ERR_INVALID, "view index out of range");
  const BView &b = hviews[view];
  ctx->make_current();
  ctx->staging.to_host(out, images.as<uint32_t>() + b.px0, (size_t)b.width * b.height * 4, ctx->stream);
}

}  // namespace dl
