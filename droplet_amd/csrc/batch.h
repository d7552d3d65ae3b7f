// Many views over several resident scenes in one pass per stage (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "binning.h"
#include "dl_internal.h"
#include "render.h"

namespace dl {

// One view of a group, as the kernels read it (device table). Camera fields
// are refreshed on every call; the rest is the group's layout.
struct BView {
  ProjUniforms u;
  double near_plane, far_plane;
  float bg[4];
  uint32_t scene;      // index into the scene table
  uint32_t n;          // items: the scene's splats
  uint32_t item0;      // first item of the view in the group's item arrays
  uint32_t blk0;       // first 256-item block (cull, rects, level 1)
  uint32_t tile_blk0;  // first rasterize block (one wave per tile)
  uint32_t row0;       // first tile row in the group's virtual image
  uint32_t width, height, tiles_x, tiles_y;
  uint64_t px0;        // first pixel of the view's image
  uint32_t bit0;       // first 64-bit word of the view's live bits (passes after the first)
};

// Views [v0, v0 + nv) rendered by one launch of every stage. Binning and
// rasterization run in front-to-back passes over each view's depth-sorted
// splats (as in the single-view renderer); pass p of the group uses the
// level-1 table at l1_off[p] with blocks[p] blocks and n1[p] counts.
struct BatchGroup {
  uint32_t v0 = 0, nv = 0;
  uint32_t items = 0, cull_blocks = 0, tiles = 0;
  uint32_t rows = 0, cols = 0, max_view_rows = 0;
  size_t seg_off = 0;      // depth-sort segment table (nv + 1 entries) in hseg
  std::vector<size_t> l1_off;
  std::vector<uint32_t> blocks;
  std::vector<size_t> n1;
  uint32_t pair_cap = 0;   // capacity of the group's tile lists (any pass)
  uint32_t bit_words = 0;  // live-bit words of all its views
};

// (tile, splat) pairs a group's lists may hold (the pair total of all its
// views, before closed tiles are skipped)
constexpr uint64_t kMaxGroupPairs = (1ull << 31);

// thrown by an eager enqueue when a group's pair total exceeds group_pairs()
struct GroupOverflow {
  size_t group;
};

struct Batch {
  Context *ctx = nullptr;
  dl_renderer_opts opts{};
  uint32_t max_cov = 16384;
  uint64_t item_cap = 1ull << 28;  // items per group
  bool use_graph = true;

  // layout of the last call (scene sizes and bands, view scenes and sizes)
  std::vector<uint64_t> layout;
  std::vector<BatchGroup> groups;
  std::vector<BView> hviews;
  std::vector<L1View> hl1;
  std::vector<SceneView> hscenes;
  std::vector<uint32_t> hseg;
  std::vector<unsigned long long> hpairs;
  uint64_t px_total = 0;
  uint32_t items_max = 0, nv_max = 0;

  DeviceBuffer d_views, d_l1, d_scenes, d_seg, d_parts, d_pairs;
  DeviceBuffer keys, vals, tkeys, tvals, rec, sort_ws, rects, pvals, ranges, images, small;
  DeviceBuffer accum, closed, open_pre;  // passes: accumulator, closed tiles, open-tile map
  // passes after the first: the open tiles' summed-area tables, packed
  // rects in item order, live bits, per-block live counts and their scan
  DeviceBuffer live_sat, live_csat, rect8, live_bits, live_cnt, live_pre, live_scan_ws;
  size_t vtiles_max = 0;
  uint32_t passes_max = 1;
  BinWorkspace bin_ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  dl_batch_stats stats{};
  // A group whose pair total exceeds group_pairs() is split at its middle
  // view (view_break: a group starts there) and the call is planned again;
  // a view that exceeds it alone is solo: the single-view renderer solo_r,
  // which cuts its own passes by pairs, renders it after the groups into
  // the same image array, outside the graph. Both marks hold for the layout
  // they were found in (break_layout). pair_budget != 0 replaces the limit
  // (to test the split at small sizes) and is the solo renderer's budget.
  uint64_t pair_budget = 0;
  uint64_t group_pairs() const { return pair_budget ? pair_budget : kMaxGroupPairs; }
  std::vector<uint64_t> break_layout;
  std::vector<uint8_t> view_break, view_solo;
  std::unique_ptr<Renderer> solo_r;

  ~Batch();
  void render(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews);
  void read_image(uint32_t view, uint8_t *out) const;

 private:
  void plan(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews);
  void run_groups(const std::vector<SceneView> &scenes, const dl_view *views, uint32_t nviews);
  void enqueue(bool eager);
  void drop_graph();
};

}  // namespace dl
