/*
 * libdroplet_hip.so -- C ABI of the MI355X (gfx950) forward splat renderer and
 * cooperative voxelizer.
 *
 * This ABI replaces the reference's WebGPU device layer and GPU kernels
 * (citations are relative to the droplet repository):
 *   src/gpu/device.ts        createDevice / uploadBuffer / readbackBuffer
 *   src/gpu/radix-sort.ts    GpuRadixSorter.sort
 *   src/render/pipeline.ts   Renderer (constructor, render, destroy)
 *   src/voxel/voxelize.ts    voxelize
 *   src/render/shaders/, src/voxel/shaders/voxelize.ts   (every WGSL kernel)
 *   src/io/ply.ts            loadPlySplats, splatBounds
 *
 * Conventions:
 *   - Every call returns DL_OK (0) or an error code; dl_last_error() gives a
 *     thread-local message for the last failure on the calling thread.
 *   - Host pointers are borrowed for the duration of the call; device memory
 *     is owned by the library. Calls are synchronous unless stated.
 *   - A dl_renderer is not re-entrant: its scratch memory is shared by its
 *     calls (the reference's Renderer has the same rule).
 *   - Structs hold plain scalars and pointers only, so bun:ffi, ctypes, JNI
 *     or cgo can bind them directly.
 */
#ifndef DROPLET_HIP_H
#define DROPLET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define DL_OK 0
#define DL_ERR_INVALID 1     /* bad argument (reference: TypeError / GPU validation error) */
#define DL_ERR_DEVICE 2      /* HIP runtime failure or no device (reference: 'No WebGPU adapter') */
#define DL_ERR_OOM 3         /* device allocation failed */
#define DL_ERR_CAPACITY 4    /* a documented size limit was exceeded */
#define DL_ERR_INTERNAL 5

#define DL_PINHOLE 0u
#define DL_EQUIRECT 1u

typedef struct dl_context dl_context;
typedef struct dl_renderer dl_renderer;
typedef struct dl_scene dl_scene;
typedef struct dl_voxel_result dl_voxel_result;
typedef struct dl_batch dl_batch;

/* SplatData (src/types.ts:23-32): struct of float32 arrays, N = count.
 * positions 3N, rotations 4N [w,x,y,z], log_scales 3N, opacity_logits N,
 * colors_dc 3N, colors_sh 3*M*N channel-major per splat (M = 0/3/8/15 for
 * sh_bands 0..3; may be NULL when sh_bands == 0). */
typedef struct {
    uint32_t count;
    const float *positions;
    const float *rotations;
    const float *log_scales;
    const float *opacity_logits;
    const float *colors_dc;
    const float *colors_sh;
    uint32_t sh_bands;
} dl_splats;

/* RenderCamera (src/types.ts:74-82). fov_y/near/far are f64 (JS numbers) so
 * the f64 cull of src/render/camera.ts:56-99 is kept. */
typedef struct {
    uint32_t width, height;
    float view_matrix[16];       /* column-major world->camera */
    double fov_y, near_plane, far_plane;
    uint32_t projection;         /* DL_PINHOLE / DL_EQUIRECT */
} dl_camera;

/* BackgroundRGBA (src/types.ts:60-65); default (0,0,0,1) when NULL is passed */
typedef struct { float r, g, b, a; } dl_background;

/* DeviceOptions (src/gpu/device.ts:11-20): only the device ordinal applies. */
typedef struct { int device_index; } dl_device_opts;

/* RendererOptions (src/render/pipeline.ts:43-48). 0 selects the reference
 * default: chunk_cap = unchunked on the device (the image is bit-identical to
 * the reference's 200000-splat chunks), max_coverage_per_splat = 16384. */
typedef struct {
    uint32_t chunk_cap;
    uint32_t max_coverage_per_splat;
    /* 0: two-level span binning straight into per-tile lists (grids up to
     * 256x1024 tiles; wider grids fall back to 1); 1: emit + stable radix
     * sort of (tile, splat) pairs. Both give identical lists and images. */
    uint32_t binning;
    /* Span binning runs in front-to-back passes over the depth-sorted
     * splats. After each pass, tiles whose pixels all have T < 1e-4 are
     * closed. Later passes bin nothing into closed tiles: the reference's
     * per-pixel loop breaks before it reads those entries, so the image is
     * unchanged. first_pass is the number of splats in the first pass,
     * and each later pass is pass_growth times larger. With both 0 the
     * renderer uses passes (max(65536, visible/8) first, growth 8) only
     * for frames with at least 2^26 pairs, and the batch always uses
     * them. full_lists = 1 bins every pair in one pass, as the reference
     * does. Debug taps always use full lists. */
    uint32_t first_pass;
    uint32_t pass_growth;
    uint32_t full_lists;
    /* digit width of the LSD radix sorts, 1..8 bits (0: 8) */
    uint32_t sort_digit_bits;
    /* Tile side in pixels: 0 or 16 (the reference's, constants.ts:19), 8
     * or 32. Other sizes change the tile lists, not the image: a frame in
     * which maxCoverage truncates a splat at 16-px tiles renders at 16 px
     * (the truncation is the reference's at 16 px), and so do frames with
     * debug taps, counting or dl_render_forward. Single-view renderer only
     * (a batch uses 16). */
    uint32_t tile_size;
    /* Rasterizer workgroup: 0 or 1, one wave per tile (the reference's
     * 16x16 workgroup per tile, rasterize.ts:36); 2, 2x2 tiles per
     * 256-thread workgroup. 32-px tiles always take one 256-thread
     * workgroup (a wave per 16-px quadrant). */
    uint32_t raster_group;
} dl_renderer_opts;

/* ---- device (replaces createDevice, src/gpu/device.ts:27-41) ---- */
int dl_context_create(const dl_device_opts *opts, dl_context **out);
void dl_context_destroy(dl_context *ctx);
const char *dl_last_error(void);
/* name of the device, e.g. "gfx950"; writes at most cap bytes */
int dl_context_device_name(dl_context *ctx, char *buf, size_t cap);

/* ---- renderer (replaces Renderer, src/render/pipeline.ts:50-592) ---- */
int dl_renderer_create(dl_context *ctx, const dl_renderer_opts *opts, dl_renderer **out);
void dl_renderer_destroy(dl_renderer *r);

/* Renderer.render (pipeline.ts:181-426): uploads the splats, renders and
 * writes W*H*4 bytes of RGBA8 (R,G,B,A byte order) into out_rgba. */
int dl_render(dl_renderer *r, const dl_splats *splats, const dl_camera *cam,
              const dl_background *bg, uint8_t *out_rgba);

/* Device-resident scene: upload once, render many frames. */
int dl_scene_create(dl_context *ctx, const dl_splats *splats, dl_scene **out);
void dl_scene_destroy(dl_scene *scene);

/* ---- PLY ingestion: loadPlySplats (src/io/ply.ts:120-239) straight into a
 * resident scene. bytes/len: the whole file (already read or mapped); the
 * header is parsed on the host with the reference's rules and messages
 * (parseHeader, ply.ts:40-91); the binary vertex section is copied to the
 * device as it lies and de-interleaved there. flip_y != 0 is the
 * reference's default (PlyLoadOptions.flipY, ply.ts:96-111). */
int dl_scene_create_from_ply(dl_context *ctx, const uint8_t *bytes, uint64_t len, int flip_y, dl_scene **out);
/* count, shBands and SH coefficients per channel held per splat: M(bands)
 * for a dl_splats scene, loadPlySplats's floor(k_rest / 3) for a PLY
 * scene (its colorsSH array, ply.ts:181-188, which the renderer reads with
 * the stride of shBands like the reference). Any pointer may be NULL. */
int dl_scene_info(const dl_scene *scene, uint32_t *count, uint32_t *sh_bands, uint32_t *sh_coeffs_per_channel);
/* The scene's SplatData arrays back on the host (loadPlySplats's result for a
 * PLY scene). Any pointer may be NULL to skip it; colors_sh takes count * 3 *
 * sh_coeffs_per_channel floats. */
int dl_scene_read(dl_scene *scene, float *positions, float *rotations, float *log_scales,
                  float *opacity_logits, float *colors_dc, float *colors_sh);
/* splatBounds (src/io/ply.ts:242-255) of the resident positions: +Infinity /
 * -Infinity per axis for an empty scene. */
int dl_scene_bounds(dl_scene *scene, double bounds_min[3], double bounds_max[3]);
/* Render a resident scene. out_rgba may be NULL: the image then stays on the
 * device (dl_renderer_read_image copies it later). */
int dl_render_scene(dl_renderer *r, const dl_scene *scene, const dl_camera *cam,
                    const dl_background *bg, uint8_t *out_rgba);
int dl_renderer_read_image(dl_renderer *r, uint8_t *out_rgba);

/* Stage timing and sizes of the last render. */
#define DL_STAGE_CULL 0
#define DL_STAGE_DEPTH_SORT 1
#define DL_STAGE_PROJECT 2
#define DL_STAGE_SCAN 3
#define DL_STAGE_EMIT 4
#define DL_STAGE_PAIR_SORT 5
#define DL_STAGE_RANGES 6
#define DL_STAGE_RASTERIZE 7
#define DL_NUM_STAGES 8
typedef struct {
    uint32_t num_visible;
    uint32_t num_chunks;
    uint32_t num_tiles;
    uint32_t tiles_x, tiles_y;
    uint32_t pair_key_bits;
    uint64_t num_pairs;          /* total over chunks */
    float stage_ms[DL_NUM_STAGES]; /* hipEvent times on the renderer's stream (timing enabled) */
    float total_ms;
    uint64_t pixel_iters;        /* rasterize loop iterations over all pixels (counting enabled) */
    uint64_t pairs_consumed;     /* tile-list entries read before each tile saturated (counting enabled) */
    uint64_t list_entries;       /* tile-list entries built (<= num_pairs; fewer when closed tiles are skipped) */
    /* span binning, counting enabled: splats binned over all passes (the live
     * ones of a compacted pass), row spans and 16-column group spans built */
    uint64_t binned_splats, row_spans, group_spans;
    /* key bits of the frame's depth sort: 32, or the span bits of a
     * speculative frame that sorted on key - (least key) */
    uint32_t depth_key_bits;
    /* tile side (px) and rasterizer group of the frame, and 1 when
     * maxCoverage truncated a splat at 16-px tiles (the frame then used 16) */
    uint32_t tile_size, raster_group, coverage_truncated;
} dl_render_stats;
int dl_renderer_set_timing(dl_renderer *r, int enabled);
/* enabled = 1: counts rasterize loop iterations (the work unit of the
 * rasterize roofline) with an instrumented kernel variant, at 16-px tiles
 * one wave per tile, and the binning tallies of dl_render_stats; 2: the
 * binning tallies only, at the frame's own tile size and rasterizer; 0:
 * off. Meant for one untimed frame. */
int dl_renderer_set_counting(dl_renderer *r, int enabled);
int dl_renderer_stats(dl_renderer *r, dl_render_stats *out);
/* A frame whose (tile, splat) pairs exceed 2^32 - 16 renders in passes of at
 * most 2^30 pairs each, cut in depth order from the per-splat tile counts
 * (the image is unchanged: passes resume like the reference's 200000-splat
 * chunks, pipeline.ts:221-423). pairs != 0 lowers both figures to `pairs`,
 * so the cut can be tested at small sizes; 0 restores the defaults. No
 * reference counterpart (the reference's chunks bound each pass by splats). */
int dl_renderer_set_pair_budget(dl_renderer *r, uint64_t pairs);
/* Rasterize loop iterations per tile row of the last frame rendered with
 * counting enabled (rasterize.ts:62-79 iterations of the row's pixels, the
 * per-row split of dl_render_stats.pixel_iters): *rows = tile rows
 * recorded (0 when the last frame did not count); out (may be NULL) takes
 * up to cap values. No reference counterpart: a parity aid. */
int dl_renderer_row_iters(dl_renderer *r, uint64_t *out, uint32_t cap, uint32_t *rows);

/* Debug taps for parity tests: when enabled, the next render keeps copies of
 * its intermediate arrays (concatenated over chunks). */
#define DL_TAP_SORTED_INDICES 0  /* u32[num_visible]  visible splat ids, front to back */
#define DL_TAP_DEPTH_KEYS 1      /* u32[num_visible]  sortable depth keys, sorted */
#define DL_TAP_PROJECTED 2       /* f32[num_visible*12] project() records in sorted order */
#define DL_TAP_TILE_COUNTS 3     /* u32[num_visible]  per-splat tile counts */
#define DL_TAP_PAIR_TILES 4      /* u32[num_pairs]    sorted tile ids */
#define DL_TAP_PAIR_SPLATS 5     /* u32[num_pairs]    chunk-local splat index */
#define DL_TAP_TILE_OFFSETS 6    /* u32[num_chunks*(num_tiles+1)] */
#define DL_TAP_ACCUM 7           /* f32[W*H*4] final (r, g, b, T) */
int dl_renderer_set_taps(dl_renderer *r, int enabled);
/* bytes: capacity of dst; returns DL_ERR_CAPACITY if too small */
int dl_renderer_tap(dl_renderer *r, int which, void *dst, size_t bytes, size_t *written);

/* ---- differentiable render (no reference counterpart: droplet renders
 * only, README.md:106). dl_render_forward renders like dl_render_scene and
 * returns the image before quantisation, out_rgb[W*H*3] = clamp(C + T*bg,
 * 0, 1) (finalize.ts:31-35). It keeps what the backward needs: the tile
 * lists, the projected records and the per-pixel (C, T). dl_render_backward
 * takes dL/d(out_rgb) and writes dL/d(every float of the scene) in the
 * dl_splats layouts. Transmittance is recomputed per tile front to back,
 * not stored per (pixel, splat). Pointers in dl_grads may be NULL. */
typedef struct {
    float *positions;            /* 3N */
    float *rotations;            /* 4N */
    float *log_scales;           /* 3N */
    float *opacity_logits;       /* N */
    float *colors_dc;            /* 3N */
    float *colors_sh;            /* 3*M*N (M = 3, 8, 15 for bands 1..3) */
} dl_grads;
int dl_render_forward(dl_renderer *r, const dl_scene *scene, const dl_camera *cam, const dl_background *bg,
                      float *out_rgb);
int dl_render_backward(dl_renderer *r, const float *grad_rgb, dl_grads *out);

/* ---- autotuner (no reference counterpart). It times candidate options on
 * the caller's own workload and keeps the fastest for later calls. The
 * candidates are the radix digit width and the front-to-back pass schedule.
 * Every candidate produces the same image. */
typedef struct {
    dl_renderer_opts chosen;     /* options in effect after the call */
    float chosen_ms;             /* wall ms per frame (per call for a batch) with them */
    float default_ms;            /* the same with the options the renderer had before */
    uint32_t candidates;         /* configurations timed */
} dl_tune_result;
int dl_renderer_autotune(dl_renderer *r, const dl_scene *scene, const dl_camera *cam, dl_tune_result *out);
/* The options dl_renderer_autotune times, without a device: stage 1 around
 * base (radix digit widths 8 / 6 / 4 and pass schedules, for scenes of
 * num_splats splats), then stage 2 around base (tile sizes 16 / 8 / 32 and
 * rasterizer workgroups; the tuner takes these around its stage-1 winner).
 * *count = the number of candidates; out (may be NULL) takes up to cap. */
int dl_renderer_tune_candidates(const dl_renderer_opts *base, uint64_t num_splats, dl_renderer_opts *out,
                                uint32_t cap, uint32_t *count);

/* ---- batched views (no reference counterpart: it stands beside
 * Renderer.render, src/render/pipeline.ts:181-185, whose per-view image it
 * reproduces bit for bit). Many views over several resident scenes go
 * through one pass of every stage. Views are grouped up to 2^28 scene
 * splats per group. The first call with a new layout (scene sizes and bands,
 * view scenes and sizes) runs directly and captures the sequence of launches
 * in a hipGraph; later calls with the same layout replace the cameras and
 * backgrounds and replay the graph. The list capacity comes from the
 * measured pair counts, and a later call that needs more re-plans by
 * itself. Views are limited to 256 x 1024 tiles (4096 x 16384 pixels). */
typedef struct {
    uint32_t scene;              /* index into the scenes array */
    dl_camera camera;
    dl_background background;
} dl_view;
typedef struct {
    uint32_t num_views;
    uint32_t num_groups;
    uint32_t graph_replays;      /* calls served by replaying the captured graph */
    uint32_t captures;           /* graphs captured so far */
    uint64_t num_pairs;          /* (tile, splat) pairs of the last call, all views */
    float device_ms;             /* device time of the last call (hipEvents around its work) */
} dl_batch_stats;
/* opts: max_coverage_per_splat as for the renderer; chunk_cap, when set,
 * is the number of scene splats per group (default 2^28). */
int dl_batch_create(dl_context *ctx, const dl_renderer_opts *opts, dl_batch **out);
void dl_batch_destroy(dl_batch *b);
/* enabled = 0: run every call directly (same kernels, no graph) */
int dl_batch_set_graph(dl_batch *b, int enabled);
int dl_batch_render(dl_batch *b, dl_scene *const *scenes, uint32_t num_scenes, const dl_view *views,
                    uint32_t num_views);
/* A group whose views hold more than 2^31 (tile, splat) pairs is split in
 * two by views, and planned again; a view that alone holds more is rendered
 * by the single-view path, which cuts its passes by pairs
 * (dl_renderer_set_pair_budget), into the same image array after the
 * groups. pairs != 0 lowers the 2^31 and is that renderer's budget, so the
 * split can be tested at small sizes; 0 restores the defaults. No reference
 * counterpart. */
int dl_batch_set_pair_budget(dl_batch *b, uint64_t pairs);
/* RGBA8 of view `view` of the last call (width*height*4 bytes) */
int dl_batch_read_image(dl_batch *b, uint32_t view, uint8_t *out_rgba);
int dl_batch_get_stats(dl_batch *b, dl_batch_stats *out);
int dl_batch_autotune(dl_batch *b, dl_scene *const *scenes, uint32_t num_scenes, const dl_view *views,
                      uint32_t num_views, dl_tune_result *out);

/* ---- stable u32 key/value sort (replaces GpuRadixSorter.sort,
 * src/gpu/radix-sort.ts:142-215, which is unstable; this one is stable).
 * Host arrays in/out; only the low key_bits bits of each key are compared. */
int dl_sort_pairs_u32(dl_context *ctx, uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t key_bits);

/* ---- the same stable sort without allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). keys and vals are DEVICE arrays of n values, sorted in place
 * on the context's stream; the call returns when they are sorted. Its
 * temporary arrays live in a fixed scratch array of the library (about 10
 * MB per GPU, held until the process ends), so n is at most
 * dl_sort_pairs_u32_noalloc_max() (2^20); larger n fails with
 * DL_ERR_CAPACITY. All calls on one GPU share that scratch: calls that
 * overlap on one GPU (several threads, or contexts on several streams) race
 * and must be serialised by the caller. */
int dl_sort_pairs_u32_noalloc(dl_context *ctx, uint32_t *d_keys, uint32_t *d_vals, uint32_t n, uint32_t key_bits);
uint64_t dl_sort_pairs_u32_noalloc_max(void);

/* ---- exclusive prefix sum of u32 (wrapping) values: the device-wide
 * replacement for the CPU scan of the sort's histogram
 * (src/gpu/radix-sort.ts:172-187), and the scan under the renderer's
 * binning and the voxelizer. Host arrays; out may alias in. The scan runs
 * `repeat` (>= 1) times back to back on the device (the same result each
 * time); *total gets the sum. */
int dl_scan_u32(dl_context *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t repeat, uint32_t *total);

/* ---- voxelizer (replaces voxelize, src/voxel/voxelize.ts:59-248) ---- */
typedef struct {
    double voxel_resolution;     /* 0 -> 0.05 */
    double opacity_cutoff;       /* <0 -> 0.1 (0 is a valid cutoff) */
    int has_bounds;              /* 1: use bounds_min/max (VoxelizeOptions.bounds) */
    double bounds_min[3], bounds_max[3];
} dl_voxel_opts;

int dl_voxelize(dl_context *ctx, const dl_splats *splats, const dl_voxel_opts *opts,
                dl_voxel_result **out);
/* number of non-empty blocks (BlockMaskBuffer.count) */
uint64_t dl_voxel_result_count(const dl_voxel_result *res);
/* blocks in the reference's insertion order (BlockMaskBuffer.blocks) */
int dl_voxel_result_blocks(const dl_voxel_result *res, uint32_t *block_idx, uint32_t *mask_lo,
                           uint32_t *mask_hi);
/* bounds and numBlocksX/Y/Z (JS numbers; -Infinity for an empty scene) */
int dl_voxel_result_bounds(const dl_voxel_result *res, double bounds_min[3], double bounds_max[3],
                           double num_blocks[3]);
/* GPU time of the voxelizer's device stages of the last call, in ms */
float dl_voxel_result_device_ms(const dl_voxel_result *res);
/* Timing and sizes of the call that made a result (no reference
 * counterpart): device_ms as above; eval_ms the per-voxel Gaussian sums
 * alone (the kernel that replaces voxel/shaders/voxelize.ts:86-147);
 * block_pairs the (block, Gaussian) pairs those sums read; grid_blocks the
 * blocks of the grid. Zero for results of the collision stages. */
typedef struct {
    float device_ms;
    float eval_ms;
    uint64_t block_pairs;
    uint64_t grid_blocks;
} dl_voxel_stats;
int dl_voxel_result_stats(const dl_voxel_result *res, dl_voxel_stats *out);
void dl_voxel_result_free(dl_voxel_result *res);

/* ---- the voxel grid in HBM and the collision stages after it.
 * The reference keeps BlockMaskBuffer on the host
 * (src/voxel/block-mask-buffer.ts:23-58) and has no code for the stages
 * after it: README.md:108 and block-mask-buffer.ts:10-11 name fill, carve
 * and mesh as grid arithmetic on the buffer. A dl_voxel_result keeps the
 * dense grid of 64-bit block masks on the device; the stages below read it
 * there and return new device-resident results. Definitions:
 *   fill:  every empty voxel that 6-connected empty voxels do not join to
 *          the grid's boundary becomes solid (enclosed cavities close);
 *   carve: every voxel that 6-connected empty voxels do not join to a voxel
 *          holding one of the points becomes solid. A point seeds the voxel
 *          floor((p - bounds_min) / voxel_resolution) (f64) when that voxel
 *          is inside the grid and empty;
 *   mesh:  one quad per maximal chain of identical maximal face runs (faces
 *          = solid voxels whose neighbour along the face's direction is
 *          empty or outside the grid); see dl_voxel_mesh_build.
 * Results of fill, carve and dl_voxel_result_create list their non-empty
 * blocks in the voxelizer's insertion order. */
typedef struct {
    const uint32_t *masks;      /* device: (lo, hi) per block at the flat block index
                                   bx + by*nb[0] + bz*nb[0]*nb[1]; voxel bit x + 4y + 16z */
    uint64_t num_blocks;        /* nb[0]*nb[1]*nb[2]; 0 for an empty grid */
    int32_t nb[3];
    double bounds_min[3];
    double voxel_resolution;
    const uint32_t *block_idx;  /* device: the non-empty blocks in insertion order */
    const uint32_t *mask_lo;
    const uint32_t *mask_hi;
    uint64_t count;
    int device;                 /* HIP device that holds the arrays */
} dl_voxel_device_view;

/* Device pointers of a result; valid until dl_voxel_result_free. */
int dl_voxel_result_device(const dl_voxel_result *res, dl_voxel_device_view *out);
/* Uploads a host BlockMaskBuffer (the fields of VoxelizeResult,
 * src/voxel/voxelize.ts:38-45): num_blocks must be integers with a product
 * below 2^32-1, block_idx below that product; repeated blocks OR-merge as
 * addBlock does (block-mask-buffer.ts:30-41). */
int dl_voxel_result_create(dl_context *ctx, const double bounds_min[3], const double bounds_max[3],
                           const double num_blocks[3], double voxel_resolution, const uint32_t *block_idx,
                           const uint32_t *mask_lo, const uint32_t *mask_hi, uint64_t count,
                           dl_voxel_result **out);
int dl_voxel_fill(dl_context *ctx, const dl_voxel_result *in, dl_voxel_result **out);
/* points: num_points (x, y, z) f64 triples, host memory */
int dl_voxel_carve(dl_context *ctx, const dl_voxel_result *in, const double *points, uint64_t num_points,
                   dl_voxel_result **out);
/* out[i] = 1 when the voxel holding points[i] is solid, else 0 (also
 * outside the grid); BlockMaskBuffer.get plus the bit test, for a batch. */
int dl_voxel_query(dl_context *ctx, const dl_voxel_result *res, const double *points, uint64_t num_points,
                   uint8_t *out);

/* Quad mesh of the solid voxels' boundary. For the normal axis a the
 * in-plane axes (u, v) are (y, z), (x, z), (x, y) for a = x, y, z. The faces
 * of one slice and row v form maximal runs [u0, u1) along u; a run starts a
 * quad unless row v-1 holds exactly the same run, and the quad covers the
 * following rows that hold exactly that run. Quads are ordered by direction
 * (+x, -x, +y, -y, +z, -z), slice, v0, u0. Quad q has vertices 4q..4q+3 at
 * (u0,v0) (u1,v0) (u1,v1) (u0,v1) on its plane, each coordinate
 * (float)(bounds_min + (double)i * voxel_resolution), and the triangles
 * 6q..6q+5 wound counter-clockwise seen from outside the solid. */
typedef struct dl_voxel_mesh dl_voxel_mesh;
int dl_voxel_mesh_build(dl_context *ctx, const dl_voxel_result *res, dl_voxel_mesh **out);
uint64_t dl_voxel_mesh_quads(const dl_voxel_mesh *m);
/* positions: 12 floats per quad; indices: 6 per quad */
int dl_voxel_mesh_read(const dl_voxel_mesh *m, float *positions, uint32_t *indices);
float dl_voxel_mesh_device_ms(const dl_voxel_mesh *m);
void dl_voxel_mesh_free(dl_voxel_mesh *m);

/* ---- numerics self-check: evaluates the device's transcendental and IEEE
 * primitives on n inputs (used by the parity tests).
 * op: 0 exp, 1 sqrt, 2 div (a/b), 3 f64->f32 depth key, 4 sigmoid,
 * 5 the rasterizer's exp for arguments in [-87, 0]. */
int dl_debug_math(dl_context *ctx, int op, const float *a, const float *b, float *out, uint32_t n);

/* ---- profiling aid: a kernel of known VALU work (k_valu_probe), blocks x
 * 256 threads, each lane 16 * iters f32 FMAs, so that rocprofv3 counter
 * runs can calibrate SQ_INSTS_VALU and its kin; *checksum (may be NULL) gets
 * the sum of the lanes' results. No reference counterpart. */
int dl_debug_valu_probe(dl_context *ctx, uint32_t blocks, uint32_t iters, double *checksum);

/* Waits until all work of the context's device is done (hipDeviceSynchronize),
 * including the renderer's second stream. */
int dl_context_synchronize(dl_context *ctx);

#ifdef __cplusplus
}
#endif
#endif
