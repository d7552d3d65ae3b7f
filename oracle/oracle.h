/*
 * CPU oracle for droplet's forward splat renderer and voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path (libdroplet_hip.so) is compared against. Only tests, the smoke check
 * and the bench's CPU-baseline leg load it; the product never links it.
 *
 * It restates, in plain C with f32/f64 arithmetic and no FMA contraction, the
 * reference's TypeScript host code and WGSL kernels (file:line citations are
 * relative to the droplet repository):
 *   camera / cull ......... src/render/camera.ts:32-99
 *   depth sort ............ src/spatial/radix-sort.ts:20-117
 *   pack .................. src/render/preprocess.ts:90-138
 *   project ............... src/render/shaders/project.ts:22-332
 *   scan .................. src/render/shaders/prefix-sum.ts:22-86, src/render/pipeline.ts:277-286
 *   emit .................. src/render/shaders/emit-pairs.ts:27-56
 *   pair sort ............. src/render/pipeline.ts:544-578
 *   tile ranges ........... src/render/shaders/find-boundaries.ts:17-99
 *   rasterize ............. src/render/shaders/rasterize.ts:36-82
 *   finalize .............. src/render/shaders/finalize.ts:29-42
 *   render orchestration .. src/render/pipeline.ts:181-426
 *   AABB .................. src/spatial/aabb.ts:22-62
 *   BVH ................... src/spatial/bvh.ts:57-177
 *   voxelize .............. src/voxel/voxelize.ts:59-306, src/voxel/shaders/voxelize.ts:17-148
 *   block mask buffer ..... src/voxel/block-mask-buffer.ts:23-58
 *
 * Transcendentals. WGSL leaves exp() precision to the adapter, so the oracle
 * fixes one function, orc_expf: Cody-Waite reduction by ln2 (two-constant
 * split) and the degree-7 Taylor polynomial of e^r on |r| <= ln2/2, evaluated
 * with fmaf (Horner), scaled by 2^k in two exact steps. It is pinned against
 * libm (<= 1 ulp). The product implements the same specification. atan2/asin
 * (equirect only) are evaluated in f64 and rounded to f32. JS f64 arithmetic
 * (cull, AABB, bounds, BVH) is restated in C double without contraction.
 *
 * Contraction. C here is compiled without contraction; WGSL lets a compiler
 * fuse `a * b + c`, and GPU back ends do. The specification fuses exactly two
 * expressions of the rasterizer's loop body (rasterize.ts:71,77), both on the
 * device and here: the Gaussian exponent
 *   power = -0.5 * fma(C*dy, dy, fma((2B)*dx, dy, (A*dx)*dx))
 * (orc_power32) and the colour accumulation color = fma(T*alpha, c, color).
 * Everything else is unfused.
 *
 * Pinning status of each part is listed in the test suite (tests/test_oracle_*.py).
 */
#ifndef DROPLET_ORACLE_H
#define DROPLET_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PINHOLE 0u
#define ORC_EQUIRECT 1u

typedef struct {
    uint32_t count;
    const float *positions;      /* 3N */
    const float *rotations;      /* 4N, [w,x,y,z] */
    const float *log_scales;     /* 3N */
    const float *opacity_logits; /* N */
    const float *colors_dc;      /* 3N */
    const float *colors_sh;      /* 3*M*N channel-major per splat, may be NULL when bands == 0 */
    uint32_t sh_bands;           /* 0..3 */
} orc_splats;

typedef struct {
    uint32_t width, height;
    float view[16];              /* column-major world->camera */
    double fov_y, near_plane, far_plane;
    uint32_t projection;         /* ORC_PINHOLE / ORC_EQUIRECT */
} orc_camera;

/* Optional outputs of orc_render. Any pointer may be NULL. Capacities are in
 * elements; when a capacity is too small nothing is written to that array and
 * the needed size is still reported. Arrays that are per chunk are the
 * concatenation over chunks in chunk order. */
typedef struct {
    /* sizes (always written) */
    uint32_t num_visible;
    uint32_t num_chunks;
    uint64_t num_pairs;          /* total over chunks */
    uint32_t num_tiles;
    /* stats (always written) */
    uint64_t pixel_iters;        /* loop iterations executed across all pixels (before the T break) */
    uint64_t pixel_evals;        /* iterations that passed the r-box test */
    uint64_t pixel_blends;       /* iterations that composited (alpha >= 1/255) */
    uint64_t tile_iters;         /* per tile: max over its pixels of iterations, summed over tiles */
    /* arrays */
    uint32_t *sorted_indices;    /* [num_visible] visible splat ids, front to back */
    uint32_t *depth_keys;        /* [num_visible] sortable f32 depth keys in sorted order */
    float *projected;            /* [num_visible * 12] project() records (3 vec4 each) */
    uint32_t *tile_counts;       /* [num_visible] per-splat tile counts (pre-scan) */
    uint32_t *pair_tiles;        /* [num_pairs] sorted tile ids */
    uint32_t *pair_splats;       /* [num_pairs] chunk-local splat index */
    uint32_t *tile_offsets;      /* [num_chunks * (num_tiles + 1)] */
    uint32_t *chunk_pairs;       /* [num_chunks] pairs in each chunk */
    float *accum;                /* [W*H*4] final (r, g, b, T) */
    uint64_t cap_visible;
    uint64_t cap_pairs;
    uint64_t cap_chunks;
} orc_render_taps;

/* --- primitives --- */
/* the rasterizer's Gaussian exponent (rasterize.ts:71), contracted as above */
static inline float orc_power32(float A, float B, float C, float dx, float dy) {
    return (-0.5f) * __builtin_fmaf(C * dy, dy, __builtin_fmaf((2.0f * B) * dx, dy, (A * dx) * dx));
}
float orc_expf(float x);
/* out[i] = fmaf(a[i], b[i], c[i]), for the tests' numpy restatements */
void orc_fmaf_array(const float *a, const float *b, const float *c, float *out, size_t n);
void orc_expf_array(const float *in, float *out, size_t n);
float orc_sigmoidf(float x);
uint32_t orc_float_to_sortable(float f);
uint32_t orc_u32sat(float f);

/* --- camera / cull / depth sort --- */
double orc_focal(const orc_camera *cam);
uint32_t orc_cull(const float *positions, uint32_t n, const orc_camera *cam,
                  uint32_t *out_indices, float *out_depths);
void orc_depth_sort(uint32_t *indices, const float *depths, uint32_t n);

/* --- project one splat from its packed record (stride 14+3M floats) ---
 * writes 12 floats into out and returns the tile count */
uint32_t orc_project_packed(const float *rec, uint32_t bands, const orc_camera *cam,
                            uint32_t max_coverage, float out[12]);

/* --- full render (pipeline.ts:181-473). chunk_cap 0 = reference default
 * (200000); max_coverage 0 = reference default (16384). bg may be NULL
 * (reference default (0,0,0,1)). out_rgba: W*H*4 bytes or NULL. */
int orc_render(const orc_splats *s, const orc_camera *cam, const float *bg,
               uint32_t chunk_cap, uint32_t max_coverage, uint8_t *out_rgba,
               orc_render_taps *taps, int num_threads);

/* Stable counting sort of (tile, value) pairs by tile (pipeline.ts:544-578). */
void orc_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n);

/* Time-bounded sample of the render: runs every CPU stage on the whole scene
 * and rasterizes only tile rows r with r % row_step == 0; returns seconds for
 * the stages in stage_seconds[0..6] = cull, depth sort, pack+project,
 * scan+emit, pair sort, ranges, rasterize(sampled rows)+finalize. */
int orc_render_timed(const orc_splats *s, const orc_camera *cam, uint32_t row_step,
                     int num_threads, double *stage_seconds, uint64_t *num_pairs);
/* The reference pipeline rasterizing only tile rows ty with ty % row_step ==
 * row_phase, and emitting only their pairs (their lists equal the full
 * emit's): out_rgba holds those rows (others show the background),
 * num_pairs counts every pair, row_iters[tiles_y] (may be NULL) gets each
 * kept row's loop iterations. -2: a chunk would hold 2^32 pairs or more. */
int orc_render_rows(const orc_splats *s, const orc_camera *cam, const float *bg, uint32_t chunk_cap,
                    uint32_t max_coverage, uint32_t row_step, uint32_t row_phase, uint8_t *out_rgba,
                    uint32_t *num_visible, uint64_t *num_pairs, uint64_t *row_iters, int num_threads);

/* --- voxelizer (voxelize.ts:59-306) --- */
typedef struct {
    double voxel_resolution;     /* default 0.05 */
    double opacity_cutoff;       /* default 0.1 */
    int has_bounds;
    double bounds_min[3], bounds_max[3];
} orc_voxel_opts;

typedef struct {
    /* always written */
    uint64_t num_blocks;         /* non-empty blocks */
    double bounds_min[3], bounds_max[3];
    double num_blocks_xyz[3];    /* JS numbers: may be -Infinity for an empty scene */
    uint64_t num_batches;        /* batches with >= 1 candidate */
    uint64_t num_candidates;     /* sum of batch candidate counts */
    uint64_t contrib_pairs;      /* (voxel, Gaussian) evaluations that passed the AABB test */
    uint64_t num_near;           /* voxels with |alpha - cutoff| <= near_eps */
    uint64_t ref_tests;          /* (voxel, candidate) AABB tests the reference runs: each voxel tests its
                                    batch's list in order until sigma >= 7 (contrib_pairs of them pass) */
    /* arrays (may be NULL; capacities in elements) */
    uint32_t *block_idx, *mask_lo, *mask_hi;   /* [num_blocks] insertion order */
    uint64_t cap_blocks;
    uint32_t *near_block, *near_bit;           /* [num_near] */
    float *near_alpha;
    uint64_t cap_near;
    double near_eps;
} orc_voxel_result;

int orc_voxelize(const orc_splats *s, const orc_voxel_opts *opts, orc_voxel_result *res,
                 int num_threads);

/* --- FP64 central-difference checker of the differentiable render
 * (gradcheck.c; pinhole). The forward is evaluated in double from the f32
 * inputs, with the f32 pipeline's lists, order and integer radii.
 * L = sum_px sum_rgb (clamp(C + T*bg, 0, 1) - target)^2. Parameters of a
 * splat are indexed 0-2 position, 3-6 rotation, 7-9 log-scale, 10 opacity
 * logit, 11-13 DC, 14.. SH (channel-major, as colors_sh). orc_gc_fd
 * returns the difference quotients plus flags: 1 not visible, 2 integer
 * radius changes, 4 depth order changes, 8 culled, 16 a per-pixel branch
 * that depends on the splat (its own test, the early stop, an output
 * clamp) differs between p+h and p-h, 32 a kink of the projection (fade,
 * eigenvalue floor, colour clamp) lies within the stencil. The f32 forward
 * decides the per-pixel branches of the checked function. */
typedef struct orc_gradcheck orc_gradcheck;
orc_gradcheck *orc_gc_create(const orc_splats *s, const orc_camera *cam, const float *bg,
                             const float *target_rgb, uint32_t max_coverage, int num_threads);
void orc_gc_free(orc_gradcheck *g);
uint32_t orc_gc_num_params(const orc_gradcheck *g);
/* f64 record of a visible splat: sx sy A B C alpha r g b, then the radius */
int orc_gc_record(const orc_gradcheck *g, uint32_t id, double out[10]);
/* f64 image (W*H*3, may be NULL) and loss */
void orc_gc_image(const orc_gradcheck *g, double *out_rgb, double *loss, int num_threads);
int orc_gc_fd(const orc_gradcheck *g, uint32_t n, const uint32_t *ids, const uint32_t *params, double h,
              double *out, int32_t *flags, int num_threads);
/* the same for one field of the splat's f64 record (0 sx, 1 sy, 2 A, 3 B,
 * 4 C, 5 alpha, 6-8 colour) */
int orc_gc_fd_record(const orc_gradcheck *g, uint32_t id, uint32_t field, double h, double *out);
void orc_aabb_extents(const orc_splats *s, float *out_ext);
/* BVH DFS leaf order (right child first) of all Gaussians, written to out_perm[N] */
int orc_bvh_order(const float *positions, const float *extents, uint32_t n, uint32_t *out_perm);

/* Morton helpers (morton.ts:11-22) */
uint32_t orc_morton_encode3(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_decode3(uint32_t m, uint32_t out[3]);

/* Collision stages (collision.c; no reference counterpart, definitions in
 * that file). `mask` is the dense grid: one 64-bit mask per block, flat
 * block index bx + by*nb0 + bz*nb0*nb1. Return 0, or -1 when out of memory. */
int orc_vox_fill(const uint64_t *mask, const int64_t nb[3], uint64_t *out);
int orc_vox_carve(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res, const double *points,
                  uint64_t npoints, uint64_t *out);
void orc_vox_query(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res,
                   const double *points, uint64_t npoints, uint8_t *out);
/* number of quads; the first cap_quads are written (12 floats, 6 indices each) */
uint64_t orc_vox_mesh(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res, float *pos,
                      uint32_t *idx, uint64_t cap_quads);

#ifdef __cplusplus
}
#endif
#endif
