/*
 * CPU oracle (test infrastructure only; see oracle.h for scope and citations).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp). Every f32
 * expression below keeps the reference's evaluation order (WGSL/JS evaluate
 * left to right) and rounds after every operation; nothing is contracted.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* WGSL prelude constants (src/render/shaders/constants.ts:19-51). Each is   */
/* the JS number printed into the shader, parsed as an abstract float and    */
/* converted to f32, hence (float)<double literal>.                          */
/* ------------------------------------------------------------------------ */
#define K_TILE 16u
#define K_SIGMA ((float)3.0)
#define K_AA ((float)0.3)
#define K_CAP ((float)0.999)
#define K_TMIN ((float)0.0001)
#define K_AMIN ((float)0.00392156862745098)
#define K_FLOOR ((float)1.1108996538e-2) /* Math.exp(-4.5).toExponential(10) */
#define K_FADE_S ((float)0.9481481481481482)
#define K_FADE_E ((float)1.8962962962962964)
#define K_DET_MIN ((float)1.0e-6)           /* project.ts:165 */
#define K_EIG_MIN ((float)0.1)              /* project.ts:181 */
#define K_TWO_PI ((float)(3.14159265 * 2.0)) /* project.ts:260 abstract const-expr */
#define K_PI ((float)3.14159265)            /* project.ts:261 */

#define CHUNK_CAP_DEFAULT 200000u   /* pipeline.ts:32 */
#define MAX_COV_DEFAULT 16384u      /* pipeline.ts:41 */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Primitives                                                               */
/* ------------------------------------------------------------------------ */
static inline float bits_to_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* e^x in f32: Cody-Waite reduction x = k ln2 + r with ln2 split into an f32
 * head and tail, degree-7 Taylor polynomial of e^r (coefficients 1/n! rounded
 * to f32) by Horner with fmaf, then 2^k applied as two exact power-of-two
 * factors so that subnormal results round once. */
float orc_expf(float x) {
    if (x != x) return x;
    if (x < -104.0f) x = -104.0f;
    if (x > 89.0f) x = 89.0f;
    float kf = rintf(x * 0x1.715476p+0f);
    float r = fmaf(kf, -0x1.62e430p-1f, x);
    r = fmaf(kf, 0x1.05c610p-29f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;
    int k1 = k / 2;
    int k2 = k - k1;
    float s1 = bits_to_f((uint32_t)(k1 + 127) << 23);
    float s2 = bits_to_f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

void orc_fmaf_array(const float *a, const float *b, const float *c, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = fmaf(a[i], b[i], c[i]);
}

void orc_expf_array(const float *in, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_expf(in[i]);
}

/* project.ts:55 */
float orc_sigmoidf(float x) { return 1.0f / (1.0f + orc_expf(-x)); }

/* radix-sort.ts:20-24 */
uint32_t orc_float_to_sortable(float f) {
    uint32_t u = f_to_bits(f);
    return (u & 0x80000000u) ? ~u : (u ^ 0x80000000u);
}

/* WGSL u32(f32): round toward zero, clamp to the u32 range (NaN -> 0). */
uint32_t orc_u32sat(float f) {
    if (!(f == f)) return 0u;
    if (f <= 0.0f) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

static inline float wmax(float a, float b) { return a > b ? a : b; }
static inline float wmin(float a, float b) { return a < b ? a : b; }
static inline float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }

/* ------------------------------------------------------------------------ */
/* Camera + cull (camera.ts:32-99)                                          */
/* ------------------------------------------------------------------------ */
double orc_focal(const orc_camera *cam) {
    return (double)cam->height / (2.0 * tan(cam->fov_y / 2.0));
}

uint32_t orc_cull(const float *pos, uint32_t n, const orc_camera *cam,
                  uint32_t *out_idx, float *out_depth) {
    const float *m = cam->view;
    uint32_t count = 0;
    if (cam->projection == ORC_PINHOLE) {
        for (uint32_t i = 0; i < n; i++) {
            double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
            double cz = (double)m[2] * x + (double)m[6] * y;
            cz = cz + (double)m[10] * z;
            cz = cz + (double)m[14];
            double depth = -cz;
            if (depth > cam->near_plane && depth < cam->far_plane) {
                out_idx[count] = i;
                out_depth[count] = (float)depth;
                count++;
            }
        }
    } else {
        double near_sq = cam->near_plane * cam->near_plane;
        for (uint32_t i = 0; i < n; i++) {
            double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
            double cx = (double)m[0] * x + (double)m[4] * y;
            cx = cx + (double)m[8] * z;
            cx = cx + (double)m[12];
            double cy = (double)m[1] * x + (double)m[5] * y;
            cy = cy + (double)m[9] * z;
            cy = cy + (double)m[13];
            double cz = (double)m[2] * x + (double)m[6] * y;
            cz = cz + (double)m[10] * z;
            cz = cz + (double)m[14];
            double r2 = cx * cx + cy * cy;
            r2 = r2 + cz * cz;
            if (r2 > near_sq) {
                out_idx[count] = i;
                out_depth[count] = (float)sqrt(r2);
                count++;
            }
        }
    }
    return count;
}

/* Stable LSD radix sort, 4 passes x 8 bits (radix-sort.ts:65-117). */
void orc_depth_sort(uint32_t *indices, const float *depths, uint32_t n) {
    if (n <= 1) return;
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *ktmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *itmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t hist[4 * 256];
    for (uint32_t i = 0; i < n; i++) keys[i] = orc_float_to_sortable(depths[i]);
    memset(hist, 0, sizeof(hist));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t k = keys[i];
        hist[k & 0xff]++;
        hist[256 + ((k >> 8) & 0xff)]++;
        hist[512 + ((k >> 16) & 0xff)]++;
        hist[768 + ((k >> 24) & 0xff)]++;
    }
    for (int p = 0; p < 4; p++) {
        uint32_t sum = 0;
        for (int b = 0; b < 256; b++) {
            uint32_t c = hist[p * 256 + b];
            hist[p * 256 + b] = sum;
            sum += c;
        }
    }
    uint32_t *in_i = indices, *in_k = keys, *out_i = itmp, *out_k = ktmp;
    for (int p = 0; p < 4; p++) {
        int shift = p * 8;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t k = in_k[i];
            uint32_t b = (k >> shift) & 0xff;
            uint32_t dst = hist[p * 256 + b]++;
            out_i[dst] = in_i[i];
            out_k[dst] = k;
        }
        uint32_t *t = in_i; in_i = out_i; out_i = t;
        t = in_k; in_k = out_k; out_k = t;
    }
    if (in_i != indices) memcpy(indices, in_i, sizeof(uint32_t) * n);
    free(keys); free(ktmp); free(itmp);
}

/* ------------------------------------------------------------------------ */
/* Project (project.ts:22-332)                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    float view[16];
    float width, height, focal, cx, cy, near_, far_;
    uint32_t tiles_x, tiles_y, max_cov, bands, proj;
} unif_t;

static uint32_t num_sh(uint32_t bands) {
    return bands == 0 ? 0u : bands == 1 ? 3u : bands == 2 ? 8u : 15u;
}

/* quatRot (project.ts:50-53): t = 2 cross(q.yzw, v); v + q.x t + cross(q.yzw, t) */
static void quat_rot(const float q[4], float v0, float v1, float v2, float o[3]) {
    float c0 = q[2] * v2 - q[3] * v1;
    float c1 = q[3] * v0 - q[1] * v2;
    float c2 = q[1] * v1 - q[2] * v0;
    float t0 = 2.0f * c0, t1 = 2.0f * c1, t2 = 2.0f * c2;
    float d0 = q[2] * t2 - q[3] * t1;
    float d1 = q[3] * t0 - q[1] * t2;
    float d2 = q[1] * t1 - q[2] * t0;
    o[0] = (v0 + q[0] * t0) + d0;
    o[1] = (v1 + q[0] * t1) + d1;
    o[2] = (v2 + q[0] * t2) + d2;
}

/* evalSH (project.ts:264-330); sh points at the packed record's offset 14 */
static void eval_sh(const float *sh, uint32_t bands, float x, float y, float z, float out[3]) {
    uint32_t M = num_sh(bands);
    float r = 0.0f, g = 0.0f, b = 0.0f;
    const float SH1 = (float)0.4886025119029199;
    float e0 = (-SH1) * y, e1 = SH1 * z, e2 = (-SH1) * x;
    r = r + e0 * sh[0]; r = r + e1 * sh[1]; r = r + e2 * sh[2];
    g = g + e0 * sh[M + 0]; g = g + e1 * sh[M + 1]; g = g + e2 * sh[M + 2];
    b = b + e0 * sh[2 * M + 0]; b = b + e1 * sh[2 * M + 1]; b = b + e2 * sh[2 * M + 2];
    if (bands >= 2) {
        float xx = x * x, yy = y * y, zz = z * z;
        float xy = x * y, yz = y * z, xz = x * z;
        float b2[5];
        b2[0] = (float)1.0925484305920792 * xy;
        b2[1] = (float)-1.0925484305920792 * yz;
        b2[2] = (float)0.31539156525252005 * ((2.0f * zz - xx) - yy);
        b2[3] = (float)-1.0925484305920792 * xz;
        b2[4] = (float)0.5462742152960396 * (xx - yy);
        for (uint32_t k = 0; k < 5; k++) {
            r = r + b2[k] * sh[3 + k];
            g = g + b2[k] * sh[M + 3 + k];
            b = b + b2[k] * sh[2 * M + 3 + k];
        }
        if (bands >= 3) {
            float b3[7];
            b3[0] = ((float)-0.5900435899266435 * y) * (3.0f * xx - yy);
            b3[1] = ((float)2.890611442640554 * xy) * z;
            b3[2] = ((float)-0.4570457994644658 * y) * ((4.0f * zz - xx) - yy);
            b3[3] = ((float)0.3731763325901154 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
            b3[4] = ((float)-0.4570457994644658 * x) * ((4.0f * zz - xx) - yy);
            b3[5] = ((float)1.445305721320277 * z) * (xx - yy);
            b3[6] = ((float)-0.5900435899266435 * x) * (xx - 3.0f * yy);
            for (uint32_t k = 0; k < 7; k++) {
                r = r + b3[k] * sh[8 + k];
                g = g + b3[k] * sh[M + 8 + k];
                b = b + b3[k] * sh[2 * M + 8 + k];
            }
        }
    }
    out[0] = r; out[1] = g; out[2] = b;
}

static uint32_t project_one(const float *in, const unif_t *u, float out[12]) {
    const float *m = u->view;
    float px = in[0], py = in[1], pz = in[2];
    float q[4] = {in[3], in[4], in[5], in[6]};
    float logit = in[10];
    float dc0 = in[11], dc1 = in[12], dc2 = in[13];
    float s0 = orc_expf(in[7]), s1 = orc_expf(in[8]), s2 = orc_expf(in[9]);

    float pcx = ((m[0] * px + m[4] * py) + m[8] * pz) + m[12] * 1.0f;
    float pcy = ((m[1] * px + m[5] * py) + m[9] * pz) + m[13] * 1.0f;
    float pcz = ((m[2] * px + m[6] * py) + m[10] * pz) + m[14] * 1.0f;
    float depth = -pcz;
    float sx, sy;

    for (int k = 0; k < 12; k++) out[k] = 0.0f;
    if (u->proj == ORC_PINHOLE) {
        if (depth <= u->near_ || depth >= u->far_) goto culled;
        float proj_x = (pcx * u->focal) / (-pcz);
        float proj_y = ((-pcy) * u->focal) / (-pcz);
        sx = proj_x + u->cx;
        sy = proj_y + u->cy;
    } else {
        float rr = sqrtf((pcx * pcx + pcy * pcy) + pcz * pcz);
        if (rr <= u->near_) goto culled;
        float lon = (float)atan2((double)pcx, (double)(-pcz));
        float lat = (float)asin((double)(pcy / rr));
        sx = (lon / K_TWO_PI + 0.5f) * u->width;
        sy = (0.5f - lat / K_PI) * u->height;
    }
    {
        float rx[3], ry[3], rz[3];
        quat_rot(q, 1.0f, 0.0f, 0.0f, rx);
        quat_rot(q, 0.0f, 1.0f, 0.0f, ry);
        quat_rot(q, 0.0f, 0.0f, 1.0f, rz);
        for (int k = 0; k < 3; k++) { rx[k] = rx[k] * s0; ry[k] = ry[k] * s1; rz[k] = rz[k] * s2; }
        float c00 = (rx[0] * rx[0] + ry[0] * ry[0]) + rz[0] * rz[0];
        float c01 = (rx[0] * rx[1] + ry[0] * ry[1]) + rz[0] * rz[1];
        float c02 = (rx[0] * rx[2] + ry[0] * ry[2]) + rz[0] * rz[2];
        float c11 = (rx[1] * rx[1] + ry[1] * ry[1]) + rz[1] * rz[1];
        float c12 = (rx[1] * rx[2] + ry[1] * ry[2]) + rz[1] * rz[2];
        float c22 = (rx[2] * rx[2] + ry[2] * ry[2]) + rz[2] * rz[2];
        float v00 = m[0], v01 = m[4], v02 = m[8];
        float v10 = m[1], v11 = m[5], v12 = m[9];
        float v20 = m[2], v21 = m[6], v22 = m[10];
        float a00 = (v00 * c00 + v01 * c01) + v02 * c02;
        float a01 = (v00 * c01 + v01 * c11) + v02 * c12;
        float a02 = (v00 * c02 + v01 * c12) + v02 * c22;
        float a10 = (v10 * c00 + v11 * c01) + v12 * c02;
        float a11 = (v10 * c01 + v11 * c11) + v12 * c12;
        float a12 = (v10 * c02 + v11 * c12) + v12 * c22;
        float a20 = (v20 * c00 + v21 * c01) + v22 * c02;
        float a21 = (v20 * c01 + v21 * c11) + v22 * c12;
        float a22 = (v20 * c02 + v21 * c12) + v22 * c22;
        float cc00 = (a00 * v00 + a01 * v01) + a02 * v02;
        float cc01 = (a00 * v10 + a01 * v11) + a02 * v12;
        float cc02 = (a00 * v20 + a01 * v21) + a02 * v22;
        float cc11 = (a10 * v10 + a11 * v11) + a12 * v12;
        float cc12 = (a10 * v20 + a11 * v21) + a12 * v22;
        float cc22 = (a20 * v20 + a21 * v21) + a22 * v22;

        float inv_z = 1.0f / pcz;
        float j00 = u->focal * inv_z;
        float j02 = (((-u->focal) * pcx) * inv_z) * inv_z;
        float j11 = u->focal * inv_z;
        float j12 = (((-u->focal) * pcy) * inv_z) * inv_z;
        float b00 = j00 * cc00 + j02 * cc02;
        float b01 = j00 * cc01 + j02 * cc12;
        float b02 = j00 * cc02 + j02 * cc22;
        float b11 = j11 * cc11 + j12 * cc12;
        float b12 = j11 * cc12 + j12 * cc22;
        float cov00 = b00 * j00 + b02 * j02;
        float cov01 = b01 * j11 + b02 * j12;
        float cov11 = b11 * j11 + b12 * j12;

        float det_orig = cov00 * cov11 - cov01 * cov01;
        cov00 = cov00 + K_AA;
        cov11 = cov11 + K_AA;
        float det_blur = cov00 * cov11 - cov01 * cov01;
        float aa = sqrtf(wmax(det_orig / det_blur, 0.0f));
        float det = det_blur;
        if (det <= K_DET_MIN) goto culled;
        float inv_det = 1.0f / det;
        float conic_a = cov11 * inv_det;
        float conic_b = (-cov01) * inv_det;
        float conic_c = cov00 * inv_det;

        float mid = 0.5f * (cov00 + cov11);
        float r_eig = sqrtf(wmax(K_EIG_MIN, mid * mid - det));
        float lambda1 = mid + r_eig;
        float radius_raw = K_SIGMA * sqrtf(lambda1);
        float fade_s = K_FADE_S * u->height;
        float fade_e = K_FADE_E * u->height;
        float fade = wclamp((fade_e - radius_raw) / (fade_e - fade_s), 0.0f, 1.0f);
        if (fade <= 0.0f) goto culled;
        float radius = ceilf(wmin(radius_raw, fade_e));

        float col[3] = {dc0, dc1, dc2};
        if (u->bands > 0) {
            float len = sqrtf((pcx * pcx + pcy * pcy) + pcz * pcz);
            float sh[3];
            eval_sh(in + 14, u->bands, pcx / len, pcy / len, pcz / len, sh);
            col[0] = col[0] + sh[0]; col[1] = col[1] + sh[1]; col[2] = col[2] + sh[2];
        }
        for (int k = 0; k < 3; k++) col[k] = wmax(col[k] + 0.5f, 0.0f);
        float alpha = (orc_sigmoidf(logit) * fade) * aa;

        out[0] = sx; out[1] = sy; out[2] = radius; out[3] = depth;
        out[4] = conic_a; out[5] = conic_b; out[6] = conic_c; out[7] = alpha;
        out[8] = col[0]; out[9] = col[1]; out[10] = col[2]; out[11] = 0.0f;

        uint32_t min_tx = orc_u32sat(wmax(0.0f, floorf((sx - radius) / 16.0f)));
        uint32_t min_ty = orc_u32sat(wmax(0.0f, floorf((sy - radius) / 16.0f)));
        uint32_t max_tx = orc_u32sat(wmin((float)u->tiles_x - 1.0f, floorf((sx + radius) / 16.0f)));
        uint32_t max_ty = orc_u32sat(wmin((float)u->tiles_y - 1.0f, floorf((sy + radius) / 16.0f)));
        uint32_t count = 0;
        if (max_tx >= min_tx && max_ty >= min_ty) {
            count = (max_tx - min_tx + 1u) * (max_ty - min_ty + 1u);
            if (count > u->max_cov) count = u->max_cov;
        }
        return count;
    }
culled:
    out[0] = 0.0f; out[1] = 0.0f; out[2] = -1.0f; out[3] = 0.0f;
    return 0u;
}

static void make_unif(const orc_camera *cam, uint32_t bands, uint32_t max_cov, unif_t *u) {
    memcpy(u->view, cam->view, sizeof(u->view));
    u->width = (float)cam->width;
    u->height = (float)cam->height;
    u->focal = (float)orc_focal(cam);
    u->cx = (float)((double)cam->width / 2.0);
    u->cy = (float)((double)cam->height / 2.0);
    u->near_ = (float)cam->near_plane;
    u->far_ = (float)cam->far_plane;
    u->tiles_x = (cam->width + K_TILE - 1) / K_TILE;
    u->tiles_y = (cam->height + K_TILE - 1) / K_TILE;
    u->max_cov = max_cov;
    u->bands = bands;
    u->proj = cam->projection;
}

uint32_t orc_project_packed(const float *rec, uint32_t bands, const orc_camera *cam,
                            uint32_t max_coverage, float out[12]) {
    unif_t u;
    make_unif(cam, bands, max_coverage ? max_coverage : MAX_COV_DEFAULT, &u);
    return project_one(rec, &u, out);
}

/* packChunkInput (preprocess.ts:90-138) for one splat */
static void pack_one(const orc_splats *s, uint32_t g, float *o) {
    uint32_t M = num_sh(s->sh_bands);
    o[0] = s->positions[3 * (size_t)g]; o[1] = s->positions[3 * (size_t)g + 1]; o[2] = s->positions[3 * (size_t)g + 2];
    for (int k = 0; k < 4; k++) o[3 + k] = s->rotations[4 * (size_t)g + k];
    for (int k = 0; k < 3; k++) o[7 + k] = s->log_scales[3 * (size_t)g + k];
    o[10] = s->opacity_logits[g];
    for (int k = 0; k < 3; k++) o[11 + k] = s->colors_dc[3 * (size_t)g + k];
    if (s->sh_bands > 0 && s->colors_sh)
        for (uint32_t k = 0; k < 3 * M; k++) o[14 + k] = s->colors_sh[(size_t)g * 3 * M + k];
}

/* ------------------------------------------------------------------------ */
/* Emit / pair sort / ranges / rasterize / finalize                         */
/* ------------------------------------------------------------------------ */
/* emitPairsShader (emit-pairs.ts:27-56) for one splat */
/* Row filter of the sampled-row mode: tile row ty is kept when
 * ty % step == phase (step 1 keeps every row). */
typedef struct { uint32_t step, phase; } rows_t;
static const rows_t ALL_ROWS = {1u, 0u};
static inline int row_kept(const rows_t *rows, uint32_t ty) { return ty % rows->step == rows->phase; }

/* emitPairsShader (emit-pairs.ts:27-56) for one splat, keeping only the pairs
 * of kept tile rows: maxCoverage counts every pair (emit-pairs.ts:48), so the
 * kept pairs are exactly those the full emit writes into kept rows, in the
 * same order. Returns the number written. With pt == NULL it only counts. */
static uint32_t emit_rows(const float *v0, uint32_t idx, uint32_t offset, const unif_t *u, const rows_t *rows,
                          uint32_t *pt, uint32_t *ps) {
    float sx = v0[0], sy = v0[1], radius = v0[2];
    if (radius <= 0.0f) return 0;
    uint32_t min_tx = orc_u32sat(wmax(0.0f, floorf((sx - radius) / 16.0f)));
    uint32_t min_ty = orc_u32sat(wmax(0.0f, floorf((sy - radius) / 16.0f)));
    uint32_t max_tx = orc_u32sat(wmin((float)u->tiles_x - 1.0f, floorf((sx + radius) / 16.0f)));
    uint32_t max_ty = orc_u32sat(wmin((float)u->tiles_y - 1.0f, floorf((sy + radius) / 16.0f)));
    if (max_tx < min_tx || max_ty < min_ty) return 0;
    uint32_t written = 0, kept = 0;
    for (uint32_t ty = min_ty; ty <= max_ty; ty++) {
        for (uint32_t tx = min_tx; tx <= max_tx; tx++) {
            if (written >= u->max_cov) return kept;
            if (row_kept(rows, ty)) {
                if (pt) {
                    pt[offset + kept] = ty * u->tiles_x + tx;
                    ps[offset + kept] = idx;
                }
                kept++;
            }
            written++;
        }
    }
    return kept;
}

static void emit_one(const float *v0, uint32_t idx, uint32_t offset, const unif_t *u,
                     uint32_t *pt, uint32_t *ps) {
    float sx = v0[0], sy = v0[1], radius = v0[2];
    if (radius <= 0.0f) return;
    uint32_t min_tx = orc_u32sat(wmax(0.0f, floorf((sx - radius) / 16.0f)));
    uint32_t min_ty = orc_u32sat(wmax(0.0f, floorf((sy - radius) / 16.0f)));
    uint32_t max_tx = orc_u32sat(wmin((float)u->tiles_x - 1.0f, floorf((sx + radius) / 16.0f)));
    uint32_t max_ty = orc_u32sat(wmin((float)u->tiles_y - 1.0f, floorf((sy + radius) / 16.0f)));
    if (max_tx < min_tx || max_ty < min_ty) return;
    uint32_t written = 0;
    for (uint32_t ty = min_ty; ty <= max_ty; ty++) {
        for (uint32_t tx = min_tx; tx <= max_tx; tx++) {
            if (written >= u->max_cov) return;
            pt[offset] = ty * u->tiles_x + tx;
            ps[offset] = idx;
            offset++;
            written++;
        }
    }
}

/* cpuSortPairs (pipeline.ts:544-578): stable counting sort by tile */
void orc_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t n) {
    if (n <= 1) return;
    uint32_t max_t = 0;
    for (uint32_t i = 0; i < n; i++) if (keys[i] > max_t) max_t = keys[i];
    uint32_t buckets = max_t + 1;
    uint32_t *counts = (uint32_t *)calloc(buckets, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) counts[keys[i]]++;
    uint32_t sum = 0;
    for (uint32_t b = 0; b < buckets; b++) { uint32_t c = counts[b]; counts[b] = sum; sum += c; }
    uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *so = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t t = keys[i];
        uint32_t dst = counts[t]++;
        to[dst] = t;
        so[dst] = vals[i];
    }
    memcpy(keys, to, sizeof(uint32_t) * n);
    memcpy(vals, so, sizeof(uint32_t) * n);
    free(to); free(so); free(counts);
}

/* init / findBoundaries / finalizeTileOffsets (find-boundaries.ts:17-99),
 * run exactly as the three kernels do. */
static void tile_ranges(const uint32_t *keys, uint32_t n, uint32_t num_tiles, uint32_t *off) {
    for (uint32_t i = 0; i <= num_tiles; i++) off[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cur = keys[i];
        if (i == 0) {
            if (off[cur] > 0u) off[cur] = 0u;
        } else {
            uint32_t prev = keys[i - 1];
            if (cur != prev)
                for (uint32_t t = prev + 1; t <= cur; t++)
                    if (off[t] > i) off[t] = i;
        }
    }
    off[num_tiles] = n;
    uint32_t fill = n;
    uint32_t i = num_tiles;
    for (;;) {
        if (i == 0) break;
        i = i - 1;
        if (off[i] == 0xFFFFFFFFu) off[i] = fill; else fill = off[i];
    }
}

typedef struct { uint64_t iters, evals, blends, tile_iters; } rstats_t;

/* rasterizeShader (rasterize.ts:36-82) for one tile */
static void raster_tile(uint32_t tx, uint32_t ty, const unif_t *u, const float *proj,
                        const uint32_t *off, const uint32_t *ps, float *accum, rstats_t *st) {
    uint32_t tile = ty * u->tiles_x + tx;
    uint32_t start = off[tile], end = off[tile + 1];
    uint32_t W = (uint32_t)u->width, H = (uint32_t)u->height;
    uint64_t tmax = 0;
    for (uint32_t ly = 0; ly < K_TILE; ly++) {
        for (uint32_t lx = 0; lx < K_TILE; lx++) {
            uint32_t x = tx * K_TILE + lx, y = ty * K_TILE + ly;
            if (x >= W || y >= H) continue;
            size_t pix = (size_t)y * W + x;
            float pxf = (float)x + 0.5f, pyf = (float)y + 0.5f;
            float cr = accum[4 * pix], cg = accum[4 * pix + 1], cb = accum[4 * pix + 2];
            float T = accum[4 * pix + 3];
            if (T < K_TMIN) continue;
            uint64_t it = 0;
            for (uint32_t i = start; i < end; i++) {
                if (T < K_TMIN) break;
                it++;
                const float *v = proj + 12 * (size_t)ps[i];
                float dx = pxf - v[0];
                float dy = pyf - v[1];
                float r = v[2];
                if (r <= 0.0f || fabsf(dx) > r || fabsf(dy) > r) continue;
                st->evals++;
                float power = orc_power32(v[4], v[5], v[6], dx, dy);
                if (power > 0.0f) continue;
                float alpha = wmin(K_CAP, v[7] * wmax(0.0f, orc_expf(power) - K_FLOOR));
                if (alpha < K_AMIN) continue;
                st->blends++;
                float w = T * alpha;
                cr = fmaf(w, v[8], cr);
                cg = fmaf(w, v[9], cg);
                cb = fmaf(w, v[10], cb);
                T = T * (1.0f - alpha);
            }
            st->iters += it;
            if (it > tmax) tmax = it;
            accum[4 * pix] = cr; accum[4 * pix + 1] = cg; accum[4 * pix + 2] = cb; accum[4 * pix + 3] = T;
        }
    }
    st->tile_iters += tmax;
}

/* rasterizes the kept tile rows; row_iters (may be NULL) gets each tile
 * row's loop iterations added */
static void rasterize(const unif_t *u, const float *proj, const uint32_t *off, const uint32_t *ps,
                      float *accum, rstats_t *st, const rows_t *rows, uint64_t *row_iters, int nthreads) {
    uint64_t it = 0, ev = 0, bl = 0, ti = 0;
    int ty_n = (int)u->tiles_y;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : it, ev, bl, ti) num_threads(nthreads)
    for (int ty = 0; ty < ty_n; ty++) {
        if (!row_kept(rows, (uint32_t)ty)) continue;
        rstats_t s = {0, 0, 0, 0};
        for (uint32_t tx = 0; tx < u->tiles_x; tx++) raster_tile(tx, (uint32_t)ty, u, proj, off, ps, accum, &s);
        if (row_iters) row_iters[ty] += s.iters;
        it += s.iters; ev += s.evals; bl += s.blends; ti += s.tile_iters;
    }
    st->iters += it; st->evals += ev; st->blends += bl; st->tile_iters += ti;
}

/* finalizeShader (finalize.ts:29-42) */
static void finalize(uint32_t W, uint32_t H, const float *accum, const float bg[4], uint8_t *out) {
    size_t npx = (size_t)W * H;
    for (size_t i = 0; i < npx; i++) {
        const float *v = accum + 4 * i;
        float T = v[3];
        float r = wclamp(v[0] + T * bg[0], 0.0f, 1.0f);
        float g = wclamp(v[1] + T * bg[1], 0.0f, 1.0f);
        float b = wclamp(v[2] + T * bg[2], 0.0f, 1.0f);
        float a = wclamp((1.0f - T) + T * bg[3], 0.0f, 1.0f);
        out[4 * i + 0] = (uint8_t)orc_u32sat(r * 255.0f + 0.5f);
        out[4 * i + 1] = (uint8_t)orc_u32sat(g * 255.0f + 0.5f);
        out[4 * i + 2] = (uint8_t)orc_u32sat(b * 255.0f + 0.5f);
        out[4 * i + 3] = (uint8_t)orc_u32sat(a * 255.0f + 0.5f);
    }
}

/* ------------------------------------------------------------------------ */
/* Render orchestration (pipeline.ts:181-473)                               */
/* ------------------------------------------------------------------------ */
/* The reference pipeline. rows selects the tile rows that are rasterized;
 * with emit_all_rows == 0 the emit keeps only the pairs of those rows too
 * (a tile's list is the same either way: the pair sort is stable and keeps
 * no pair of another tile between two of its own), so a frame whose full
 * lists would not fit in memory can be checked on a sample of its rows.
 * num_pairs always counts every pair. Returns -2 when a chunk would hold
 * 2^32 or more pairs (the u32 pair index of the reference's buffers). */
static int render_impl(const orc_splats *s, const orc_camera *cam, const float *bg_in,
                       uint32_t chunk_cap_opt, uint32_t max_cov, uint8_t *out_rgba,
                       orc_render_taps *taps, int nthreads, const rows_t *rows, int emit_all_rows,
                       uint64_t *row_iters, double *stage_s) {
    if (!s || !cam) return -1;
    if (s->sh_bands > 3) return -1;
    float bg[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (bg_in) memcpy(bg, bg_in, sizeof(bg));
    if (max_cov == 0) max_cov = MAX_COV_DEFAULT;
    if (chunk_cap_opt == 0) chunk_cap_opt = CHUNK_CAP_DEFAULT;
    uint32_t bands = s->sh_bands;
    uint32_t M = num_sh(bands);
    uint32_t stride = 14 + 3 * M;
    unif_t u;
    make_unif(cam, bands, max_cov, &u);
    uint32_t num_tiles = u.tiles_x * u.tiles_y;
    size_t npx = (size_t)cam->width * cam->height;
    double t0 = now_s(), t1;

    uint32_t n = s->count;
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    float *dep = (float *)malloc(sizeof(float) * (n ? n : 1));
    uint32_t nvis = orc_cull(s->positions, n, cam, idx, dep);
    t1 = now_s(); if (stage_s) stage_s[0] += t1 - t0; t0 = t1;
    uint32_t *keys_unsorted = NULL;
    if (taps && taps->depth_keys) {
        keys_unsorted = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        /* keys in cull order so the sorted keys can be reported by id */
        for (uint32_t i = 0; i < nvis; i++) keys_unsorted[idx[i]] = orc_float_to_sortable(dep[i]);
    }
    orc_depth_sort(idx, dep, nvis);
    t1 = now_s(); if (stage_s) stage_s[1] += t1 - t0; t0 = t1;

    float *accum = (float *)malloc(sizeof(float) * 4 * (npx ? npx : 1));
    for (size_t p = 0; p < npx; p++) { accum[4 * p] = 0; accum[4 * p + 1] = 0; accum[4 * p + 2] = 0; accum[4 * p + 3] = 1.0f; }

    rstats_t st = {0, 0, 0, 0};
    uint64_t pairs_total = 0, pairs_written = 0;
    uint32_t nchunks = 0;
    int rc = 0;
    if (nvis > 0) {
        uint32_t chunk_cap = chunk_cap_opt < nvis ? chunk_cap_opt : nvis;
        float *stage = (float *)malloc(sizeof(float) * (size_t)stride * chunk_cap);
        float *proj = (float *)malloc(sizeof(float) * 12 * (size_t)chunk_cap);
        uint32_t *counts = (uint32_t *)malloc(sizeof(uint32_t) * chunk_cap);
        uint32_t *offs = (uint32_t *)malloc(sizeof(uint32_t) * chunk_cap);
        uint32_t *toff = (uint32_t *)malloc(sizeof(uint32_t) * (num_tiles + 1));
        for (uint32_t c0 = 0; c0 < nvis; c0 += chunk_cap) {
            uint32_t cn = (nvis - c0) < chunk_cap ? (nvis - c0) : chunk_cap;
            uint32_t chunk_id = nchunks++;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t i = 0; i < (int64_t)cn; i++) {
                pack_one(s, idx[c0 + i], stage + (size_t)i * stride);
                counts[i] = project_one(stage + (size_t)i * stride, &u, proj + 12 * (size_t)i);
            }
            t1 = now_s(); if (stage_s) stage_s[2] += t1 - t0; t0 = t1;
            if (taps) {
                if (taps->projected && (uint64_t)c0 + cn <= taps->cap_visible)
                    memcpy(taps->projected + 12 * (size_t)c0, proj, sizeof(float) * 12 * cn);
                if (taps->tile_counts && (uint64_t)c0 + cn <= taps->cap_visible)
                    memcpy(taps->tile_counts + c0, counts, sizeof(uint32_t) * cn);
            }
            uint64_t sum = 0;
            for (uint32_t i = 0; i < cn; i++) sum += counts[i];
            if (sum > 0xFFFFFFFFull) { rc = -2; free(stage); free(proj); free(counts); free(offs); free(toff); goto done; }
            pairs_total += sum;
            if (taps && taps->chunk_pairs && chunk_id < taps->cap_chunks) taps->chunk_pairs[chunk_id] = (uint32_t)sum;
            if (!emit_all_rows) {
                /* counts of the kept rows only */
#pragma omp parallel for schedule(static) num_threads(nthreads)
                for (int64_t i = 0; i < (int64_t)cn; i++)
                    counts[i] = emit_rows(proj + 12 * (size_t)i, (uint32_t)i, 0, &u, rows, NULL, NULL);
                sum = 0;
                for (uint32_t i = 0; i < cn; i++) sum += counts[i];
            }
            {
                uint64_t run = 0;
                for (uint32_t i = 0; i < cn; i++) { offs[i] = (uint32_t)run; run += counts[i]; }
            }
            uint32_t npairs = (uint32_t)sum;
            if (npairs == 0) {
                if (taps && taps->tile_offsets && chunk_id < taps->cap_chunks)
                    for (uint32_t t = 0; t <= num_tiles; t++) taps->tile_offsets[(size_t)chunk_id * (num_tiles + 1) + t] = 0;
                continue;
            }
            uint32_t *pt = (uint32_t *)malloc(sizeof(uint32_t) * npairs);
            uint32_t *ps = (uint32_t *)malloc(sizeof(uint32_t) * npairs);
            if (emit_all_rows) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
                for (int64_t i = 0; i < (int64_t)cn; i++) emit_one(proj + 12 * (size_t)i, (uint32_t)i, offs[i], &u, pt, ps);
            } else {
#pragma omp parallel for schedule(static) num_threads(nthreads)
                for (int64_t i = 0; i < (int64_t)cn; i++)
                    (void)emit_rows(proj + 12 * (size_t)i, (uint32_t)i, offs[i], &u, rows, pt, ps);
            }
            t1 = now_s(); if (stage_s) stage_s[3] += t1 - t0; t0 = t1;
            orc_sort_pairs(pt, ps, npairs);
            t1 = now_s(); if (stage_s) stage_s[4] += t1 - t0; t0 = t1;
            tile_ranges(pt, npairs, num_tiles, toff);
            t1 = now_s(); if (stage_s) stage_s[5] += t1 - t0; t0 = t1;
            if (taps) {
                if (taps->pair_tiles && pairs_written + npairs <= taps->cap_pairs) {
                    memcpy(taps->pair_tiles + pairs_written, pt, sizeof(uint32_t) * npairs);
                    if (taps->pair_splats) memcpy(taps->pair_splats + pairs_written, ps, sizeof(uint32_t) * npairs);
                }
                if (taps->tile_offsets && chunk_id < taps->cap_chunks)
                    memcpy(taps->tile_offsets + (size_t)chunk_id * (num_tiles + 1), toff, sizeof(uint32_t) * (num_tiles + 1));
            }
            pairs_written += npairs;
            rasterize(&u, proj, toff, ps, accum, &st, rows, row_iters, nthreads);
            t1 = now_s(); if (stage_s) stage_s[6] += t1 - t0; t0 = t1;
            free(pt); free(ps);
        }
        free(stage); free(proj); free(counts); free(offs); free(toff);
    }
    if (out_rgba) finalize(cam->width, cam->height, accum, bg, out_rgba);
    t1 = now_s(); if (stage_s) stage_s[6] += t1 - t0; t0 = t1;
    if (taps) {
        taps->num_visible = nvis;
        taps->num_chunks = nchunks;
        taps->num_pairs = pairs_total;
        taps->num_tiles = num_tiles;
        taps->pixel_iters = st.iters;
        taps->pixel_evals = st.evals;
        taps->pixel_blends = st.blends;
        taps->tile_iters = st.tile_iters;
        if (taps->sorted_indices && nvis <= taps->cap_visible) memcpy(taps->sorted_indices, idx, sizeof(uint32_t) * nvis);
        if (taps->depth_keys && keys_unsorted && nvis <= taps->cap_visible)
            for (uint32_t j = 0; j < nvis; j++) taps->depth_keys[j] = keys_unsorted[idx[j]];
        if (taps->accum) memcpy(taps->accum, accum, sizeof(float) * 4 * npx);
    }
done:
    free(keys_unsorted);
    free(idx); free(dep); free(accum);
    return rc;
}

int orc_render(const orc_splats *s, const orc_camera *cam, const float *bg, uint32_t chunk_cap,
               uint32_t max_coverage, uint8_t *out_rgba, orc_render_taps *taps, int num_threads) {
    if (num_threads <= 0) num_threads = 1;
    return render_impl(s, cam, bg, chunk_cap, max_coverage, out_rgba, taps, num_threads, &ALL_ROWS, 1, NULL, NULL);
}

int orc_render_rows(const orc_splats *s, const orc_camera *cam, const float *bg, uint32_t chunk_cap,
                    uint32_t max_coverage, uint32_t row_step, uint32_t row_phase, uint8_t *out_rgba,
                    uint32_t *num_visible, uint64_t *num_pairs, uint64_t *row_iters, int num_threads) {
    if (num_threads <= 0) num_threads = 1;
    if (row_step == 0) return -1;
    const rows_t rows = {row_step, row_phase % row_step};
    orc_render_taps taps;
    memset(&taps, 0, sizeof(taps));
    int rc = render_impl(s, cam, bg, chunk_cap, max_coverage, out_rgba, &taps, num_threads, &rows, 0, row_iters, NULL);
    if (num_visible) *num_visible = taps.num_visible;
    if (num_pairs) *num_pairs = taps.num_pairs;
    return rc;
}

int orc_render_timed(const orc_splats *s, const orc_camera *cam, uint32_t row_step, int num_threads,
                     double *stage_seconds, uint64_t *num_pairs) {
    if (num_threads <= 0) num_threads = 1;
    if (row_step == 0) row_step = 1;
    orc_render_taps taps;
    memset(&taps, 0, sizeof(taps));
    for (int k = 0; k < 7; k++) stage_seconds[k] = 0.0;
    /* one chunk, as the device path runs; every pair is emitted and sorted,
     * only every row_step-th tile row is rasterized */
    const rows_t rows = {row_step, 0u};
    int rc = render_impl(s, cam, NULL, 0xFFFFFFFFu, 0, NULL, &taps, num_threads, &rows, 1, NULL, stage_seconds);
    if (num_pairs) *num_pairs = taps.num_pairs;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Voxelizer                                                                */
/* ------------------------------------------------------------------------ */
/* computeAabbExtents (aabb.ts:22-62), f64 */
void orc_aabb_extents(const orc_splats *s, float *out) {
    for (uint32_t i = 0; i < s->count; i++) {
        double qw = s->rotations[4 * (size_t)i], qx = s->rotations[4 * (size_t)i + 1];
        double qy = s->rotations[4 * (size_t)i + 2], qz = s->rotations[4 * (size_t)i + 3];
        double sx = exp((double)s->log_scales[3 * (size_t)i]);
        double sy = exp((double)s->log_scales[3 * (size_t)i + 1]);
        double sz = exp((double)s->log_scales[3 * (size_t)i + 2]);
        double xx = qx * qx, yy = qy * qy, zz = qz * qz;
        double xy = qx * qy, xz = qx * qz, yz = qy * qz;
        double wx = qw * qx, wy = qw * qy, wz = qw * qz;
        double r00 = 1 - 2 * (yy + zz), r01 = 2 * (xy - wz), r02 = 2 * (xz + wy);
        double r10 = 2 * (xy + wz), r11 = 1 - 2 * (xx + zz), r12 = 2 * (yz - wx);
        double r20 = 2 * (xz - wy), r21 = 2 * (yz + wx), r22 = 1 - 2 * (xx + yy);
        double a, b, c;
        a = r00 * sx; b = r01 * sy; c = r02 * sz;
        double ex = sqrt((a * a + b * b) + c * c) * 3.0;
        a = r10 * sx; b = r11 * sy; c = r12 * sz;
        double ey = sqrt((a * a + b * b) + c * c) * 3.0;
        a = r20 * sx; b = r21 * sy; c = r22 * sz;
        double ez = sqrt((a * a + b * b) + c * c) * 3.0;
        out[3 * (size_t)i] = (float)ex;
        out[3 * (size_t)i + 1] = (float)ey;
        out[3 * (size_t)i + 2] = (float)ez;
    }
}

/* GaussianBVH (bvh.ts:33-183) */
typedef struct {
    double mn[3], mx[3];
    uint32_t start, end;
    int32_t left, right;
} bvh_node;

typedef struct {
    bvh_node *nodes;
    size_t n_nodes, cap_nodes;
    uint32_t *idx;
    const float *c, *e;
} bvh_t;

static void bvh_nth(bvh_t *b, uint32_t lo, uint32_t hi, uint32_t k, int axis) {
    uint32_t *idx = b->idx;
    const float *c = b->c;
    while (hi - lo > 1) {
        uint32_t m = (lo + hi) >> 1;
        float a = c[(size_t)idx[lo] * 3 + axis];
        float bb = c[(size_t)idx[m] * 3 + axis];
        float d = c[(size_t)idx[hi - 1] * 3 + axis];
        uint32_t pivot_idx = (a < bb) ? (bb < d ? m : (a < d ? hi - 1 : lo))
                                      : (a < d ? lo : (bb < d ? hi - 1 : m));
        uint32_t t0 = idx[lo + 1]; idx[lo + 1] = idx[pivot_idx]; idx[pivot_idx] = t0;
        float pivot = c[(size_t)idx[lo + 1] * 3 + axis];
        int64_t i = lo + 1, j = hi - 1;
        for (;;) {
            while (++i < (int64_t)hi && c[(size_t)idx[i] * 3 + axis] < pivot) {}
            while (--j > (int64_t)lo && c[(size_t)idx[j] * 3 + axis] > pivot) {}
            if (i >= j) break;
            uint32_t t = idx[i]; idx[i] = idx[j]; idx[j] = t;
        }
        uint32_t t2 = idx[lo + 1]; idx[lo + 1] = idx[j]; idx[j] = t2;
        if ((uint32_t)j == k) return;
        if ((uint32_t)j < k) lo = (uint32_t)j + 1; else hi = (uint32_t)j;
    }
}

static int32_t bvh_build(bvh_t *b, uint32_t start, uint32_t end) {
    if (b->n_nodes == b->cap_nodes) {
        b->cap_nodes = b->cap_nodes ? 2 * b->cap_nodes : 1024;
        b->nodes = (bvh_node *)realloc(b->nodes, sizeof(bvh_node) * b->cap_nodes);
    }
    int32_t ni = (int32_t)b->n_nodes++;
    bvh_node node;
    for (int k = 0; k < 3; k++) { node.mn[k] = INFINITY; node.mx[k] = -INFINITY; }
    node.start = start; node.end = end; node.left = -1; node.right = -1;
    for (uint32_t i = start; i < end; i++) {
        uint32_t g = b->idx[i];
        for (int k = 0; k < 3; k++) {
            double cc = b->c[(size_t)g * 3 + k], ee = b->e[(size_t)g * 3 + k];
            if (cc - ee < node.mn[k]) node.mn[k] = cc - ee;
            if (cc + ee > node.mx[k]) node.mx[k] = cc + ee;
        }
    }
    b->nodes[ni] = node;
    uint32_t count = end - start;
    if (count <= 8) return ni;
    double dx = node.mx[0] - node.mn[0], dy = node.mx[1] - node.mn[1], dz = node.mx[2] - node.mn[2];
    int axis = dx > dy ? (dx > dz ? 0 : 2) : (dy > dz ? 1 : 2);
    uint32_t mid = start + (count >> 1);
    bvh_nth(b, start, end, mid, axis);
    int32_t l = bvh_build(b, start, mid);
    int32_t r = bvh_build(b, mid, end);
    b->nodes[ni].left = l;
    b->nodes[ni].right = r;
    return ni;
}

static void bvh_init(bvh_t *b, const float *pos, const float *ext, uint32_t n) {
    memset(b, 0, sizeof(*b));
    b->idx = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) b->idx[i] = i;
    b->c = pos; b->e = ext;
    if (n > 0) bvh_build(b, 0, n);
}

static void bvh_free(bvh_t *b) { free(b->nodes); free(b->idx); }

/* queryOverlappingInto (bvh.ts:136-177). out_len plays out.length. */
static int64_t bvh_query(const bvh_t *b, const double qmn[3], const double qmx[3],
                         uint32_t *out, uint64_t out_off, uint64_t out_len) {
    if (b->n_nodes == 0) return 0;
    int64_t count = 0;
    int32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const bvh_node *nd = &b->nodes[stack[--sp]];
        if (nd->mx[0] < qmn[0] || nd->mn[0] > qmx[0]) continue;
        if (nd->mx[1] < qmn[1] || nd->mn[1] > qmx[1]) continue;
        if (nd->mx[2] < qmn[2] || nd->mn[2] > qmx[2]) continue;
        if (nd->left < 0) {
            for (uint32_t i = nd->start; i < nd->end; i++) {
                uint32_t g = b->idx[i];
                int ok = 1;
                for (int k = 0; k < 3 && ok; k++) {
                    double cc = b->c[(size_t)g * 3 + k], ee = b->e[(size_t)g * 3 + k];
                    double gmn = cc - ee, gmx = cc + ee;
                    if (gmx < qmn[k] || gmn > qmx[k]) ok = 0;
                }
                if (!ok) continue;
                if (out_off + (uint64_t)count >= out_len) return -1;
                out[out_off + (uint64_t)count++] = g;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return count;
}

int orc_bvh_order(const float *positions, const float *extents, uint32_t n, uint32_t *out_perm) {
    bvh_t b;
    bvh_init(&b, positions, extents, n);
    double qmn[3] = {-INFINITY, -INFINITY, -INFINITY}, qmx[3] = {INFINITY, INFINITY, INFINITY};
    int64_t c = bvh_query(&b, qmn, qmx, out_perm, 0, n);
    bvh_free(&b);
    return c == (int64_t)n ? 0 : -1;
}

uint32_t orc_morton_encode3(uint32_t x, uint32_t y, uint32_t z) {
    return ((x & 1) << 0) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) | ((z & 2) << 4);
}

void orc_morton_decode3(uint32_t m, uint32_t o[3]) {
    o[0] = (m & 1) | ((m >> 2) & 2);
    o[1] = ((m >> 1) & 1) | ((m >> 3) & 2);
    o[2] = ((m >> 2) & 1) | ((m >> 4) & 2);
}

/* JS Math.round: nearest integer, ties toward +Infinity */
static double js_round(double x) {
    if (!isfinite(x)) return x;
    double f = floor(x);
    return (x - f >= 0.5) ? f + 1.0 : f;
}

/* Gaussian record as uploaded (voxelize.ts:250-276) */
typedef struct { float p[3], logit, q[4], ls[3], ext[3]; } vgauss_t;

/* evalGaussian (voxel/shaders/voxelize.ts:62-84) */
static float eval_gauss(const float vc[3], float half, const vgauss_t *g, int *passed) {
    float diff[3];
    for (int k = 0; k < 3; k++) diff[k] = vc[k] - g->p[k];
    for (int k = 0; k < 3; k++)
        if (fabsf(diff[k]) > g->ext[k] + half) return 0.0f;
    *passed = 1;
    float cd[3];
    for (int k = 0; k < 3; k++) {
        float cl = wmin(wmax(g->p[k], vc[k] - half), vc[k] + half);
        cd[k] = cl - g->p[k];
    }
    float qx = -g->q[1], qy = -g->q[2], qz = -g->q[3], qw = g->q[0];
    float t0 = 2.0f * (qy * cd[2] - qz * cd[1]);
    float t1 = 2.0f * (qz * cd[0] - qx * cd[2]);
    float t2 = 2.0f * (qx * cd[1] - qy * cd[0]);
    float l0 = (cd[0] + qw * t0) + (qy * t2 - qz * t1);
    float l1 = (cd[1] + qw * t1) + (qz * t0 - qx * t2);
    float l2 = (cd[2] + qw * t2) + (qx * t1 - qy * t0);
    float i0 = orc_expf(-g->ls[0]), i1 = orc_expf(-g->ls[1]), i2 = orc_expf(-g->ls[2]);
    float s0 = l0 * i0, s1 = l1 * i1, s2 = l2 * i2;
    float d2 = (s0 * s0 + s1 * s1) + s2 * s2;
    float op = 1.0f / (1.0f + orc_expf(-g->logit));
    return op * orc_expf((-0.5f) * d2);
}

typedef struct { uint32_t bx, by, bz, cx, cy, cz; double mn[3]; uint64_t off, cnt; } vbatch_t;

int orc_voxelize(const orc_splats *s, const orc_voxel_opts *opts, orc_voxel_result *res, int nthreads) {
    if (nthreads <= 0) nthreads = 1;
    double vres = opts ? opts->voxel_resolution : 0.05;
    double cutoff = opts ? opts->opacity_cutoff : 0.1;
    double bs = 4.0 * vres;
    uint32_t n = s->count;
    float *ext = (float *)malloc(sizeof(float) * 3 * (n ? n : 1));
    orc_aabb_extents(s, ext);
    bvh_t bvh;
    bvh_init(&bvh, s->positions, ext, n);

    double bmn[3], bmx[3];
    if (opts && opts->has_bounds) {
        for (int k = 0; k < 3; k++) { bmn[k] = opts->bounds_min[k]; bmx[k] = opts->bounds_max[k]; }
    } else {
        /* computeBlockAlignedBounds (voxelize.ts:278-306) */
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                double p = s->positions[3 * (size_t)i + k], e = ext[3 * (size_t)i + k];
                if (p - e < mn[k]) mn[k] = p - e;
                if (p + e > mx[k]) mx[k] = p + e;
            }
        for (int k = 0; k < 3; k++) { bmn[k] = floor(mn[k] / bs) * bs; bmx[k] = ceil(mx[k] / bs) * bs; }
    }
    double nbd[3];
    for (int k = 0; k < 3; k++) nbd[k] = js_round((bmx[k] - bmn[k]) / bs);
    for (int k = 0; k < 3; k++) { res->bounds_min[k] = bmn[k]; res->bounds_max[k] = bmx[k]; res->num_blocks_xyz[k] = nbd[k]; }
    res->num_blocks = 0; res->num_batches = 0; res->num_candidates = 0; res->contrib_pairs = 0; res->num_near = 0;
    res->ref_tests = 0;

    vgauss_t *gs = (vgauss_t *)malloc(sizeof(vgauss_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) gs[i].p[k] = s->positions[3 * (size_t)i + k];
        gs[i].logit = s->opacity_logits[i];
        for (int k = 0; k < 4; k++) gs[i].q[k] = s->rotations[4 * (size_t)i + k];
        for (int k = 0; k < 3; k++) gs[i].ls[k] = s->log_scales[3 * (size_t)i + k];
        for (int k = 0; k < 3; k++) gs[i].ext[k] = ext[3 * (size_t)i + k];
    }

    int64_t nbx = (nbd[0] > 0 && isfinite(nbd[0])) ? (int64_t)nbd[0] : 0;
    int64_t nby = (nbd[1] > 0 && isfinite(nbd[1])) ? (int64_t)nbd[1] : 0;
    int64_t nbz = (nbd[2] > 0 && isfinite(nbd[2])) ? (int64_t)nbd[2] : 0;
    const uint64_t MEGA = 2u * 1024u * 1024u;
    uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * MEGA);
    float resf = (float)vres, half = resf * 0.5f, cutf = (float)cutoff;
    int rc = 0;
    uint64_t contrib_total = 0, tests_total = 0;

    for (int64_t bz = 0; bz < nbz && rc == 0; bz += 16)
        for (int64_t by = 0; by < nby && rc == 0; by += 16)
            for (int64_t bx = 0; bx < nbx && rc == 0; bx += 16) {
                vbatch_t bt;
                bt.bx = (uint32_t)bx; bt.by = (uint32_t)by; bt.bz = (uint32_t)bz;
                bt.cx = (uint32_t)(nbx - bx < 16 ? nbx - bx : 16);
                bt.cy = (uint32_t)(nby - by < 16 ? nby - by : 16);
                bt.cz = (uint32_t)(nbz - bz < 16 ? nbz - bz : 16);
                double qmn[3], qmx[3];
                qmn[0] = bmn[0] + (double)bx * bs; qmn[1] = bmn[1] + (double)by * bs; qmn[2] = bmn[2] + (double)bz * bs;
                qmx[0] = qmn[0] + (double)bt.cx * bs; qmx[1] = qmn[1] + (double)bt.cy * bs; qmx[2] = qmn[2] + (double)bt.cz * bs;
                /* every batch is evaluated on its own list, so the flush
                 * grouping never changes a result; a query that does not fit
                 * the 2M index buffer even at offset 0 is the reference's
                 * unhandled case (voxelize.ts:204-225) and is an error here */
                int64_t cnt = bvh_query(&bvh, qmn, qmx, scratch, 0, MEGA);
                if (cnt < 0) { rc = -2; break; }
                if (cnt == 0) continue;
                res->num_batches++;
                res->num_candidates += (uint64_t)cnt;
                float bminf[3] = {(float)qmn[0], (float)qmn[1], (float)qmn[2]};
                uint32_t total = bt.cx * bt.cy * bt.cz;
                uint32_t *lo_arr = (uint32_t *)calloc(total, sizeof(uint32_t));
                uint32_t *hi_arr = (uint32_t *)calloc(total, sizeof(uint32_t));
                uint64_t contrib = 0, tests = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : contrib, tests)
                {
                    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cnt);
                    /* position of each kept candidate in the batch list */
                    uint32_t *cpos = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cnt);
#pragma omp for schedule(dynamic, 4)
                    for (int64_t fb = 0; fb < (int64_t)total; fb++) {
                        uint32_t blk_x = (uint32_t)fb % bt.cx;
                        uint32_t blk_y = ((uint32_t)fb / bt.cx) % bt.cy;
                        uint32_t blk_z = (uint32_t)fb / (bt.cx * bt.cy);
                        float bo[3] = {((float)blk_x * 4.0f) * resf, ((float)blk_y * 4.0f) * resf, ((float)blk_z * 4.0f) * resf};
                        /* Exact block-level prefilter: a Gaussian that fails
                         * the per-voxel AABB test for every voxel of the block
                         * adds exactly +0.0 to every sigma, so dropping it
                         * changes no sum and no early-out. The f64 window
                         * below is wider than any voxel centre by a margin
                         * far above f32 rounding. */
                        double lo_c[3], hi_c[3];
                        for (int k = 0; k < 3; k++) {
                            double base = (double)bminf[k] + (double)bo[k];
                            lo_c[k] = base + 0.5 * (double)resf;
                            hi_c[k] = base + 3.5 * (double)resf;
                        }
                        uint32_t nc = 0;
                        for (int64_t c = 0; c < cnt; c++) {
                            const vgauss_t *g = &gs[scratch[c]];
                            int keep = 1;
                            for (int k = 0; k < 3 && keep; k++) {
                                double reach = (double)g->ext[k] + (double)half;
                                double margin = 1e-5 * (fabs(lo_c[k]) + fabs((double)g->p[k]) + reach + 1.0);
                                if ((double)g->p[k] < lo_c[k] - reach - margin || (double)g->p[k] > hi_c[k] + reach + margin) keep = 0;
                            }
                            if (keep) { cpos[nc] = (uint32_t)c; cand[nc++] = scratch[c]; }
                        }
                        uint32_t mlo = 0, mhi = 0;
                        for (uint32_t vi = 0; vi < 64; vi++) {
                            uint32_t lp[3];
                            orc_morton_decode3(vi, lp);
                            float vc[3];
                            for (int k = 0; k < 3; k++) vc[k] = (bminf[k] + bo[k]) + (((float)lp[k] + 0.5f) * resf);
                            float sigma = 0.0f;
                            /* the reference tests every candidate of the
                             * batch in list order until sigma >= 7
                             * (voxel/shaders/voxelize.ts:115-131); the
                             * dropped ones fail its AABB test */
                            uint64_t voxel_tests = (uint64_t)cnt;
                            for (uint32_t c = 0; c < nc; c++) {
                                int passed = 0;
                                sigma = sigma + eval_gauss(vc, half, &gs[cand[c]], &passed);
                                contrib += (uint64_t)passed;
                                if (sigma >= 7.0f) { voxel_tests = (uint64_t)cpos[c] + 1; break; }
                            }
                            tests += voxel_tests;
                            float op = 1.0f - orc_expf(-sigma);
                            uint32_t lin = lp[2] * 16u + lp[1] * 4u + lp[0];
                            if (op >= cutf) {
                                if (lin < 32) mlo |= 1u << lin; else mhi |= 1u << (lin - 32);
                            }
                            if (res->near_eps > 0.0 && fabs((double)op - (double)cutf) <= res->near_eps) {
                                uint64_t slot;
#pragma omp atomic capture
                                slot = res->num_near++;
                                if (res->near_block && slot < res->cap_near) {
                                    res->near_block[slot] = (uint32_t)((int64_t)(bx + blk_x) + (int64_t)(by + blk_y) * nbx + (int64_t)(bz + blk_z) * nbx * nby);
                                    res->near_bit[slot] = lin;
                                    if (res->near_alpha) res->near_alpha[slot] = op;
                                }
                            }
                        }
                        lo_arr[fb] = mlo; hi_arr[fb] = mhi;
                    }
                    free(cand);
                    free(cpos);
                }
                contrib_total += contrib;
                tests_total += tests;
                /* mask decode + addBlock (voxelize.ts:160-179) */
                for (uint32_t fb = 0; fb < total; fb++) {
                    if (lo_arr[fb] == 0 && hi_arr[fb] == 0) continue;
                    uint32_t lx = fb % bt.cx, ly = (fb / bt.cx) % bt.cy, lz = fb / (bt.cx * bt.cy);
                    int64_t flat = (bx + lx) + (by + ly) * nbx + (bz + lz) * nbx * nby;
                    uint64_t k = res->num_blocks++;
                    if (res->block_idx && k < res->cap_blocks) {
                        res->block_idx[k] = (uint32_t)flat;
                        res->mask_lo[k] = lo_arr[fb];
                        res->mask_hi[k] = hi_arr[fb];
                    }
                }
                free(lo_arr); free(hi_arr);
            }
    res->contrib_pairs = contrib_total;
    res->ref_tests = tests_total;
    free(scratch); free(gs); free(ext);
    bvh_free(&bvh);
    return rc;
}
