// Renderer state and kernel parameter blocks (host + device visible).
#pragma once

#include <algorithm>
#include <array>
#include <vector>

#include "binning.h"
#include "dl_internal.h"

namespace dl {

// Floats per splat of a scene's colour record: colorsDC (3) then the 3 * M
// SH coefficients the renderer reads (channel-major, as colorsSH holds them),
// zero-padded to 16 B, 64 B, 128 B and 192 B for bands 0-3, so that the
// gather of one splat's colour data touches at most two 128-byte lines.
__host__ __device__ constexpr uint32_t color_record_floats(uint32_t bands) {
  return bands == 0 ? 4u : bands == 1 ? 16u : bands == 2 ? 32u : 48u;
}

// Device-resident SplatData (src/types.ts:23-32), same array layouts, plus
// the colour records built from colorsDC / colorsSH (color_record_floats).
struct SceneView {
  const float *pos;
  const float *rot;
  const float *ls;
  const float *op;
  const float *dc;
  const float *sh;
  const float *col;
  uint32_t count;
  uint32_t bands;
};

// Project uniforms (pipeline.ts:475-506): every value as the f32 the reference
// writes into its uniform buffer.
// tg is the tile grid the tile counts and rects use (16-px tiles as in the
// reference, or the autotuner's 8 or 32; tiles_x / tiles_y are its size);
// max_cov is maxCoverage in those tiles; max_cov16 is maxCoverage itself,
// against which every splat's 16-px range is checked so that a frame at
// another tile size can tell whether the reference would truncate a splat.
struct ProjUniforms {
  float view[16];
  float width, height, focal, cx, cy, near_, far_;
  uint32_t tiles_x, tiles_y, max_cov, proj;
  TileGrid tg;
  uint32_t max_cov16;
};

ProjUniforms make_uniforms(const dl_camera &cam, uint32_t max_cov, uint32_t tile = kTile);
// maxCoverage in tiles of `tile` pixels that truncates no splat whose 16-px
// rect holds at most max_cov tiles (8-px rects are at most 4x as large;
// 32-px rects are never larger than the 16-px ones)
inline uint32_t max_cov_for_tile(uint32_t max_cov, uint32_t tile) {
  return tile == 8 ? (uint32_t)std::min<uint64_t>(4ull * max_cov, 0xFFFFFFFFull) : max_cov;
}

// Default schedule of the front-to-back passes: a first pass of
// max(kPassMinFirst, visible / kPassFirstDiv) splats, each later pass
// kPassGrowth times larger; the single-view renderer uses passes by default
// from kPassMinPairs (tile, splat) pairs per frame (measured on MI355X:
// below that the ~30 launches of a pass cost more than the binning saved).
constexpr uint64_t kPassMinPairs = 64ull << 20;
constexpr uint64_t kPassMinFirst = 65536;
constexpr uint64_t kPassFirstDiv = 8;
constexpr uint64_t kPassGrowth = 8;
// tiles from which the first pass launches its tiles in last frame's work
// order (the ordering sort costs about four launches)
constexpr uint32_t kOrderMinTiles = 4096;

struct RastParams {
  uint32_t tiles_x, width, height;
  uint32_t has_pairs, resume, finalize, write_accum;
  // with closed tracking: a tile that closes writes its RGBA8 at once and no
  // accumulator, and later passes (the final one too) skip it
  uint32_t final_on_close;
  float bg_r, bg_g, bg_b, bg_a;
  uint8_t *closed;  // per (virtual) tile: 1 once every pixel has T < 1e-4 (null: not tracked)
  // launch order of the tiles (null: tile = workgroup index), and where each
  // tile records the sort key of its work (null: not recorded)
  const uint32_t *order;
  uint32_t *cost_key;
  // counting frames: loop iterations per tile row (null: not recorded)
  unsigned long long *row_iters;
  // tile rows, the 16-px grid (the 32-px rasterizer's quadrant filter) and
  // the width of the grid whose lists off indexes (32-px frames, where
  // tiles_x is the 16-px width)
  uint32_t tiles_y, t16x, t16y, list_tiles_x;
};

// True while the scene with this serial number exists (scenes register
// themselves on construction and leave on destruction).
bool scene_alive(uint64_t serial);

struct Scene {
  Scene();
  ~Scene();
  Scene(const Scene &) = delete;
  Scene &operator=(const Scene &) = delete;
  // unique per scene in the process, never reused
  const uint64_t serial;
  Context *ctx = nullptr;
  uint32_t count = 0, bands = 0;
  // SH coefficients per channel held per splat: M(bands) for uploaded
  // SplatData; loadPlySplats's floor(k_rest / 3) for a PLY scene, whose
  // colorsSH array the renderer reads with stride 3 * M like the reference
  uint32_t sh_coeffs = 0;
  DeviceBuffer pos, rot, ls, op, dc, sh;
  DeviceBuffer col;  // colour records (color_record_floats(bands) per splat)
  void upload(const dl_splats &s);
  // builds col from dc and sh (after upload / load_ply); waits for the stream
  void build_color_records();
  // loadPlySplats (ply.ts:120-239) into this scene (ply.hip)
  void load_ply(const uint8_t *bytes, size_t len, bool flip_y);
  // splatBounds (ply.ts:242-255) of the resident positions
  void bounds(double lo[3], double hi[3]);
  SceneView view() const;
};

struct Renderer {
  Context *ctx = nullptr;
  dl_renderer_opts opts{};
  uint32_t max_cov = 16384;
  Scene own_scene;
  // per splat (id order): depth keys/ids, projected records, tile counts
  DeviceBuffer keys, ids, tkeys, tids, sort_ws, rec, counts, parts;
  // per splat (id order): packed tile rect for span binning's later passes
  DeviceBuffer rect8;
  // per chunk
  DeviceBuffer chunk_counts, offsets, scan_ws, ranges;
  // pass cuts above the pair budget: count prefix, block sums, cut list
  DeviceBuffer cut_ws;
  // pairs
  DeviceBuffer pkeys, pvals, ptkeys, ptvals, psort_ws;
  // span binning
  BinWorkspace bin_ws;
  // image; closed-tile flags and open-tile map of front-to-back passes
  DeviceBuffer accum, rgba, small, closed, open_pre;
  // counting frames: rasterize loop iterations per tile row (u64[tiles_y])
  DeviceBuffer row_iters;
  uint32_t row_iters_rows = 0;
  uint32_t width = 0, height = 0;
  dl_render_stats stats{};
  bool timing = false;
  bool keep_taps = false;
  bool counting = false;
  bool tallying = false;  // binning tallies only (dl_renderer_set_counting 2)
  // Frames of span binning normally run without a host read of the
  // visible-splat and pair totals after the depth sort: passes cover all
  // sorted splats (invisible ones have no tiles) and the pair capacity is
  // the largest total seen so far plus headroom. The totals are read at the
  // end of the frame; a frame that needed more capacity is rendered again
  // with the exact totals. spec_cap is 0 until a frame has run.
  uint64_t spec_cap = 0;
  uint64_t last_pairs_all = 0;
  // Bits of the visible depth keys' span (max - min + 1, plus headroom)
  // seen in the last frame: a speculative frame sorts on the keys less their
  // minimum with that many bits, in fewer passes, and is rendered again with
  // the full 32-bit sort when its own span did not fit. 0 until a frame ran.
  uint32_t spec_key_bits = 0;
  uint32_t key_bits_used = 0;  // bits of the current frame's mapped sort (0: raw 32-bit keys)
  // extra bits over the last span: 0, raised after each frame whose span
  // outgrew the bits (up to 2), so a growing range stops re-rendering
  uint32_t key_headroom = 0;
  bool force_exact = false;
  // (tile, splat) pair totals above budget() in one frame do not fit the
  // u32 list offsets (2^32 - 16): such a frame is cut into passes of at
  // most cut_pairs() pairs each, 2^30 (about 20 GB of lists and spans per
  // pass), as the reference renders any total in 200000-splat chunks. On
  // the 6M-splat 4096x2048 equirect frame 2^30 measured fastest: 21 passes
  // in 5.56 ms, against 6.4 ms at 2^29 (42 passes of fixed launches) and
  // 6.7 ms at 2^31 (11 passes binning twice the list entries).
  // pair_budget != 0 sets both, to test the cut at small sizes.
  uint64_t pair_budget = 0;
  uint64_t budget() const { return pair_budget ? pair_budget : 0xFFFFFFF0ull; }
  uint64_t cut_pairs() const { return pair_budget ? pair_budget : (1ull << 30); }
  // Launch order of the first pass's tiles in large frames: decreasing work
  // of the same tiles in the last frame, in bands of about 1.4x (a 6-bit key
  // the rasterizer writes), and tile order inside a band, which keeps the
  // concurrently running tiles close together for L2. The heaviest tiles
  // then start first and the launch does not end on a few late heavy ones.
  // The order changes no result: the tiles are independent.
  // The sort that turns one frame's keys into the next frame's order runs on
  // a second stream (default priority) while pass 2 bins, and the next
  // frame's first pass waits for it (order_done).
  // With 2x2 groups or 32-px tiles the order and keys are per launch unit.
  DeviceBuffer tile_key, tile_key_tmp, tile_order, tile_order_tmp, tile_sort_ws;
  uint32_t tile_key_tiles = 0;  // launch units of the recorded keys (0: none)
  uint32_t tile_key_shape = 0;  // their tile size * 4 + group
  // Tile size and rasterizer group (opts.tile_size, opts.raster_group).
  // A frame at 8 or 32 px that finds a splat truncated by maxCoverage at
  // 16-px tiles renders again at 16 px (force_tile16), and last_cov16 starts
  // the following frames at 16 px until a frame finds none.
  bool force_tile16 = false, last_cov16 = false;
  hipStream_t order_stream = nullptr;
  hipEvent_t keys_done = nullptr, order_done = nullptr;
  const uint32_t *next_order = nullptr;  // set once an ordering sort was launched
  std::array<std::vector<uint8_t>, 8> taps;
  std::vector<uint32_t> tap_rank;  // splat id -> depth rank (taps only)
  std::vector<hipEvent_t> events;
  // pinned host block of the frame's small readbacks (the totals read after
  // the frame, an exact frame's visible and pair totals): the runtime
  // returns a readback into pageable memory about 30 us later
  static constexpr size_t kHostSmallBytes = 256;
  void *host_small = nullptr;
  void *pinned_small();
  std::vector<int> ev_stage;
  size_t ev_used = 0;

  // differentiable render (backward.hip): the forward keeps one pass of full
  // lists, the accumulator and the projected records for the backward
  // train_valid holds from a forward until the next render of any kind;
  // train_scene is the serial of the forward's scene, which the backward
  // checks is still alive (its arrays are read through train_sc)
  bool train_mode = false, train_valid = false;
  uint64_t train_scene = 0;
  SceneView train_sc{};
  ProjUniforms train_u{};
  dl_background train_bg{};
  const uint32_t *train_vals = nullptr;
  DeviceBuffer img_rgb, grad_img, gsplat, gparams;

  ~Renderer();
  void render(const SceneView &sc, const dl_camera &cam, const dl_background &bg);
  void forward_train(const SceneView &sc, uint64_t scene_serial, const dl_camera &cam, const dl_background &bg,
                     float *out_rgb);
  void backward(const float *grad_rgb, const dl_grads &out);
  int digit_bits() const { return opts.sort_digit_bits ? (int)std::min<uint32_t>(opts.sort_digit_bits, 8) : 8; }
  void read_image(uint8_t *out);

 private:
  void rerender_at16(const SceneView &sc, const dl_camera &cam, const dl_background &bg);
  void mark(int stage);
  void ensure_events(size_t n);
};

}  // namespace dl
