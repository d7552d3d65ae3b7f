/*
 * FP64 central-difference checker for the differentiable render.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). The device backward differentiates
 * the forward of the renderer (project.ts:59-234 then rasterize.ts:62-79,
 * finalize.ts:31-35 before quantisation). This file evaluates the same
 * function in double precision from the same f32 inputs. It takes the f32
 * pipeline's tile lists, depth order and integer radii from orc_render: they
 * are piecewise constant and carry no gradient. L = sum over pixels and RGB
 * of (clamp(C + T bg, 0, 1) - target)^2. A central difference (L(p+h) -
 * L(p-h)) / 2h on one parameter needs only the pixels of the tiles whose
 * lists hold that splat.
 *
 * Threshold crossings are reported as flags instead of numbers:
 *   - an integer radius that moves;
 *   - a depth order that changes;
 *   - a visibility change;
 *   - a piecewise part of the projection (fade, eigenvalue floor, colour
 *     clamp) whose side differs between p+h and p-h.
 *   - a per-pixel branch that depends on the perturbed splat (its own test,
 *     the early stop, an output clamp) and whose f64 outcome differs
 *     between p+h and p-h.
 * The f32 forward runs in lockstep with the f64 values and decides every
 * per-pixel branch (pixel_loss_forced). The other splats of a pixel enter
 * with their f32 records, exactly as the forward computed them: they are
 * constants of the function being differenced. The perturbed splat enters
 * as its f32 record plus the exact f64 change of its projection at p +- h.
 * So the quotients differentiate the function the device evaluates: the
 * projection's derivative in f64, and the compositing around the f32
 * values. (For very elongated splats the f32 conic is ill-conditioned at
 * pixels far along the long axis, and exp(power) there can differ from
 * the exact-math value by about 1%.)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GC_TILE 16u

/* the f32 constants of the forward, as doubles */
static const double kAA_ = (double)(float)0.3;
static const double kCap = (double)(float)0.999;
static const double kMinT_ = (double)(float)0.0001;
static const double kMinA = (double)(float)0.00392156862745098;
static const double kFloor = (double)(float)1.1108996538e-2;
static const double kFadeS = (double)(float)0.9481481481481482;
static const double kFadeE = (double)(float)1.8962962962962964;
static const double kDetMin_ = (double)(float)1.0e-6;
static const double kEigMin_ = (double)(float)0.1;
static const double kSigma_ = 3.0;

struct orc_gradcheck {
    orc_splats s;
    orc_camera cam;
    double bg[3];
    const float *target;         /* W*H*3 */
    uint32_t W, H, tiles_x, tiles_y, bands, M, nparam;
    double view[16], focal, cx, cy, near_, far_, height;
    uint32_t nvis;
    uint32_t *order;             /* [nvis] splat ids front to back */
    uint32_t *rank;              /* [count] rank or UINT32_MAX */
    uint32_t *keys;              /* [nvis] f32 sortable depth keys in order */
    double *rec;                 /* [nvis * 9] sx sy A B C a r g b */
    double *radius;              /* [nvis] integer radius of the f32 forward (<= 0: none) */
    float *proj32;               /* [nvis * 12] the f32 forward's records, depth order */
    uint32_t *toff;              /* [num_tiles + 1] */
    uint32_t *tlist;             /* [pairs] ranks, depth order inside each tile */
    uint32_t *stoff;             /* [nvis + 1] CSR: tiles of each rank */
    uint32_t *stiles;
};

static uint32_t gc_num_sh(uint32_t bands) { return bands == 0 ? 0u : bands == 1 ? 3u : bands == 2 ? 8u : 15u; }

static void gc_params(const orc_gradcheck *g, uint32_t id, double *p) {
    const orc_splats *s = &g->s;
    for (int k = 0; k < 3; k++) p[k] = s->positions[3 * (size_t)id + k];
    for (int k = 0; k < 4; k++) p[3 + k] = s->rotations[4 * (size_t)id + k];
    for (int k = 0; k < 3; k++) p[7 + k] = s->log_scales[3 * (size_t)id + k];
    p[10] = s->opacity_logits[id];
    for (int k = 0; k < 3; k++) p[11 + k] = s->colors_dc[3 * (size_t)id + k];
    for (uint32_t k = 0; k < 3 * g->M; k++) p[14 + k] = s->colors_sh[(size_t)id * 3 * g->M + k];
}

static void qrot(const double q[4], const double v[3], double o[3]) {
    double c0 = q[2] * v[2] - q[3] * v[1], c1 = q[3] * v[0] - q[1] * v[2], c2 = q[1] * v[1] - q[2] * v[0];
    double t0 = 2 * c0, t1 = 2 * c1, t2 = 2 * c2;
    double d0 = q[2] * t2 - q[3] * t1, d1 = q[3] * t0 - q[1] * t2, d2 = q[1] * t1 - q[2] * t0;
    o[0] = v[0] + q[0] * t0 + d0;
    o[1] = v[1] + q[0] * t1 + d1;
    o[2] = v[2] + q[0] * t2 + d2;
}

static void sh_eval64(const double *c, uint32_t M, double x, double y, double z, double out[3]) {
    double b[15];
    const double SH1 = (double)(float)0.4886025119029199;
    b[0] = -SH1 * y; b[1] = SH1 * z; b[2] = -SH1 * x;
    if (M >= 8) {
        double xx = x * x, yy = y * y, zz = z * z;
        b[3] = (double)(float)1.0925484305920792 * x * y;
        b[4] = (double)(float)-1.0925484305920792 * y * z;
        b[5] = (double)(float)0.31539156525252005 * (2 * zz - xx - yy);
        b[6] = (double)(float)-1.0925484305920792 * x * z;
        b[7] = (double)(float)0.5462742152960396 * (xx - yy);
        if (M >= 15) {
            b[8] = (double)(float)-0.5900435899266435 * y * (3 * xx - yy);
            b[9] = (double)(float)2.890611442640554 * x * y * z;
            b[10] = (double)(float)-0.4570457994644658 * y * (4 * zz - xx - yy);
            b[11] = (double)(float)0.3731763325901154 * z * (2 * zz - 3 * xx - 3 * yy);
            b[12] = (double)(float)-0.4570457994644658 * x * (4 * zz - xx - yy);
            b[13] = (double)(float)1.445305721320277 * z * (xx - yy);
            b[14] = (double)(float)-0.5900435899266435 * x * (xx - 3 * yy);
        }
    }
    for (int ch = 0; ch < 3; ch++) {
        double acc = 0;
        for (uint32_t k = 0; k < M; k++) acc += b[k] * c[ch * M + k];
        out[ch] = acc;
    }
}

/* f64 projection of one splat (pinhole). Returns 0 when culled; writes the
 * 9 differentiable values, the integer radius and the depth. */
static int proj64(const orc_gradcheck *g, const double *p, double out[9], double *radius, double *depth,
                  uint32_t *state) {
    const double *m = g->view;
    double pcx = m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12];
    double pcy = m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13];
    double pcz = m[2] * p[0] + m[6] * p[1] + m[10] * p[2] + m[14];
    *depth = -pcz;
    if (-pcz <= g->near_ || -pcz >= g->far_) return 0;
    double f = g->focal;
    double sx = (pcx * f) / (-pcz) + g->cx;
    double sy = ((-pcy) * f) / (-pcz) + g->cy;
    double q[4] = {p[3], p[4], p[5], p[6]};
    double s[3] = {exp(p[7]), exp(p[8]), exp(p[9])};
    double Mx[3][3];
    for (int j = 0; j < 3; j++) {
        double e[3] = {0, 0, 0}, o[3];
        e[j] = 1.0;
        qrot(q, e, o);
        for (int a = 0; a < 3; a++) Mx[a][j] = o[a] * s[j];
    }
    double C3[3][3], W[3][3] = {{m[0], m[4], m[8]}, {m[1], m[5], m[9]}, {m[2], m[6], m[10]}}, V[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C3[a][b] = Mx[a][0] * Mx[b][0] + Mx[a][1] * Mx[b][1] + Mx[a][2] * Mx[b][2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double acc = 0;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) acc += W[a][k] * C3[k][l] * W[b][l];
            V[a][b] = acc;
        }
    double iz = 1.0 / pcz;
    double j00 = f * iz, j02 = -f * pcx * iz * iz, j11 = f * iz, j12 = -f * pcy * iz * iz;
    double c00 = j00 * j00 * V[0][0] + 2 * j00 * j02 * V[0][2] + j02 * j02 * V[2][2];
    double c01 = j00 * j11 * V[0][1] + j00 * j12 * V[0][2] + j02 * j11 * V[1][2] + j02 * j12 * V[2][2];
    double c11 = j11 * j11 * V[1][1] + 2 * j11 * j12 * V[1][2] + j12 * j12 * V[2][2];
    double det_orig = c00 * c11 - c01 * c01;
    double d00 = c00 + kAA_, d11 = c11 + kAA_;
    double det = d00 * d11 - c01 * c01;
    double ratio = det_orig / det;
    double aa = sqrt(ratio > 0 ? ratio : 0);
    if (det <= kDetMin_) return 0;
    double mid = 0.5 * (d00 + d11);
    double disc = mid * mid - det;
    double r_eig = sqrt(disc > kEigMin_ ? disc : kEigMin_);
    double rr = kSigma_ * sqrt(mid + r_eig);
    double fs = kFadeS * g->height, fe = kFadeE * g->height;
    double fade = (fe - rr) / (fe - fs);
    /* piecewise parts of the projection: which side of each kink */
    uint32_t st = (fade < 1 ? 1u : 0u) | (disc > kEigMin_ ? 2u : 0u) | (ratio > 0 ? 4u : 0u);
    fade = fade < 0 ? 0 : fade > 1 ? 1 : fade;
    if (fade <= 0) return 0;
    *radius = ceil(rr < fe ? rr : fe);
    double col[3] = {p[11], p[12], p[13]};
    if (g->M) {
        double len = sqrt(pcx * pcx + pcy * pcy + pcz * pcz), sh[3];
        sh_eval64(p + 14, g->M, pcx / len, pcy / len, pcz / len, sh);
        for (int c = 0; c < 3; c++) col[c] += sh[c];
    }
    double sig = 1.0 / (1.0 + exp(-p[10]));
    out[0] = sx; out[1] = sy;
    out[2] = d11 / det; out[3] = -c01 / det; out[4] = d00 / det;
    out[5] = sig * fade * aa;
    for (int c = 0; c < 3; c++) {
        out[6 + c] = col[c] + 0.5 > 0 ? col[c] + 0.5 : 0.0;
        if (col[c] + 0.5 > 0) st |= 8u << c;
    }
    if (state) *state = st;
    return 1;
}

/* per-pixel compositing in f64 over one tile list; `sub_rank` takes the
 * record `sub` instead of its own. Returns the loss of the pixel and writes
 * a signature of its branches (for crossing checks) into *sig. */
static double pixel_loss(const orc_gradcheck *g, uint32_t x, uint32_t y, uint32_t t, uint32_t sub_rank,
                         const double *sub, double sub_radius, uint64_t *sig, double out_rgb[3]) {
    double T = 1.0, C[3] = {0, 0, 0};
    double px = x + 0.5, py = y + 0.5;
    uint64_t h = 1469598103934665603ull;
    uint32_t steps = 0;
    for (uint32_t e = g->toff[t]; e < g->toff[t + 1]; e++) {
        if (T < kMinT_) break;
        steps++;
        uint32_t k = g->tlist[e];
        const double *r = k == sub_rank ? sub : g->rec + 9 * (size_t)k;
        double rad = k == sub_rank ? sub_radius : g->radius[k];
        double dx = px - r[0], dy = py - r[1];
        int code;
        if (rad <= 0 || fabs(dx) > rad || fabs(dy) > rad) code = 1;
        else {
            double power = -0.5 * (r[2] * dx * dx + 2 * r[3] * dx * dy + r[4] * dy * dy);
            if (power > 0) code = 2;
            else {
                double ge = exp(power) - kFloor;
                double raw = r[5] * (ge > 0 ? ge : 0);
                double a = raw < kCap ? raw : kCap;
                if (a < kMinA) code = 3;
                else {
                    code = raw < kCap ? 4 : 5;
                    for (int c = 0; c < 3; c++) C[c] += T * a * r[6 + c];
                    T *= 1.0 - a;
                }
            }
        }
        if (k == sub_rank) h = (h ^ (uint64_t)(code + 8 * steps)) * 1099511628211ull;
    }
    double loss = 0;
    for (int c = 0; c < 3; c++) {
        double o = C[c] + T * g->bg[c];
        int cl = o <= 0 ? 0 : o >= 1 ? 2 : 1;
        h = (h ^ (uint64_t)(16 + cl)) * 1099511628211ull;
        o = o < 0 ? 0 : o > 1 ? 1 : o;
        if (out_rgb) out_rgb[c] = o;
        double d = o - (double)g->target[3 * ((size_t)y * g->W + x) + c];
        loss += d * d;
    }
    h = (h ^ (uint64_t)(1000 + steps)) * 1099511628211ull;
    *sig = h;
    return loss;
}

orc_gradcheck *orc_gc_create(const orc_splats *s, const orc_camera *cam, const float *bg, const float *target_rgb,
                             uint32_t max_coverage, int nthreads) {
    if (!s || !cam || !target_rgb || cam->projection != ORC_PINHOLE || s->sh_bands > 3) return NULL;
    orc_gradcheck *g = (orc_gradcheck *)calloc(1, sizeof(*g));
    g->s = *s;
    g->cam = *cam;
    g->target = target_rgb;
    for (int c = 0; c < 3; c++) g->bg[c] = bg ? (double)bg[c] : 0.0;
    g->W = cam->width;
    g->H = cam->height;
    g->tiles_x = (g->W + GC_TILE - 1) / GC_TILE;
    g->tiles_y = (g->H + GC_TILE - 1) / GC_TILE;
    g->bands = s->sh_bands;
    g->M = gc_num_sh(s->sh_bands);
    g->nparam = 14 + 3 * g->M;
    for (int k = 0; k < 16; k++) g->view[k] = cam->view[k];
    g->focal = (double)(float)orc_focal(cam);
    g->cx = (double)(float)((double)cam->width / 2.0);
    g->cy = (double)(float)((double)cam->height / 2.0);
    g->near_ = (double)(float)cam->near_plane;
    g->far_ = (double)(float)cam->far_plane;
    g->height = (double)(float)cam->height;
    /* lists of the f32 pipeline in one chunk */
    orc_render_taps tp;
    memset(&tp, 0, sizeof(tp));
    orc_render(s, cam, bg, 0xFFFFFFFFu, max_coverage, NULL, &tp, nthreads);
    uint32_t num_tiles = tp.num_tiles;
    g->nvis = tp.num_visible;
    uint64_t np = tp.num_pairs;
    g->order = (uint32_t *)malloc(sizeof(uint32_t) * (g->nvis + 1));
    g->keys = (uint32_t *)malloc(sizeof(uint32_t) * (g->nvis + 1));
    float *proj = (float *)malloc(sizeof(float) * 12 * ((size_t)g->nvis + 1));
    uint32_t *pt = (uint32_t *)malloc(sizeof(uint32_t) * (np + 1));
    g->tlist = (uint32_t *)malloc(sizeof(uint32_t) * (np + 1));
    g->toff = (uint32_t *)malloc(sizeof(uint32_t) * (num_tiles + 1));
    memset(&tp, 0, sizeof(tp));
    tp.sorted_indices = g->order;
    tp.depth_keys = g->keys;
    tp.projected = proj;
    tp.pair_tiles = pt;
    tp.pair_splats = g->tlist;
    tp.tile_offsets = g->toff;
    tp.cap_visible = g->nvis;
    tp.cap_pairs = np;
    tp.cap_chunks = 1;
    orc_render(s, cam, bg, 0xFFFFFFFFu, max_coverage, NULL, &tp, nthreads);
    if (g->nvis == 0 || np == 0) {
        for (uint32_t t = 0; t <= num_tiles; t++) g->toff[t] = 0;
    }
    g->rank = (uint32_t *)malloc(sizeof(uint32_t) * (s->count + 1));
    for (uint32_t i = 0; i < s->count; i++) g->rank[i] = 0xFFFFFFFFu;
    for (uint32_t j = 0; j < g->nvis; j++) g->rank[g->order[j]] = j;
    g->rec = (double *)calloc((size_t)9 * (g->nvis + 1), sizeof(double));
    g->radius = (double *)malloc(sizeof(double) * (g->nvis + 1));
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t j = 0; j < (int64_t)g->nvis; j++) {
        double p[14 + 45], rad = 0, dep = 0;
        gc_params(g, g->order[j], p);
        if (!proj64(g, p, g->rec + 9 * (size_t)j, &rad, &dep, NULL)) memset(g->rec + 9 * (size_t)j, 0, 9 * sizeof(double));
        g->radius[j] = proj[12 * (size_t)j + 2];  /* the f32 forward's integer radius (or -1) */
    }
    /* tiles of each rank */
    g->stoff = (uint32_t *)calloc((size_t)g->nvis + 2, sizeof(uint32_t));
    for (uint64_t e = 0; e < np; e++) g->stoff[g->tlist[e] + 1]++;
    for (uint32_t j = 0; j < g->nvis; j++) g->stoff[j + 1] += g->stoff[j];
    g->stiles = (uint32_t *)malloc(sizeof(uint32_t) * (np + 1));
    uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * (g->nvis + 1));
    memcpy(fill, g->stoff, sizeof(uint32_t) * (g->nvis + 1));
    for (uint32_t t = 0; t < num_tiles; t++)
        for (uint32_t e = g->toff[t]; e < g->toff[t + 1]; e++) g->stiles[fill[g->tlist[e]]++] = t;
    free(fill);
    g->proj32 = proj;
    free(pt);
    return g;
}

void orc_gc_free(orc_gradcheck *g) {
    if (!g) return;
    free(g->order); free(g->rank); free(g->keys); free(g->rec); free(g->radius);
    free(g->toff); free(g->tlist); free(g->stoff); free(g->stiles); free(g->proj32);
    free(g);
}

uint32_t orc_gc_num_params(const orc_gradcheck *g) { return g ? g->nparam : 0; }

int orc_gc_record(const orc_gradcheck *g, uint32_t id, double out[10]) {
    if (!g || id >= g->s.count || g->rank[id] == 0xFFFFFFFFu) return -1;
    uint32_t k = g->rank[id];
    memcpy(out, g->rec + 9 * (size_t)k, 9 * sizeof(double));
    out[9] = g->radius[k];
    return 0;
}

void orc_gc_image(const orc_gradcheck *g, double *out_rgb, double *loss, int nthreads) {
    double total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t y = 0; y < (int64_t)g->H; y++)
        for (uint32_t x = 0; x < g->W; x++) {
            uint64_t sig;
            uint32_t t = (uint32_t)(y / GC_TILE) * g->tiles_x + x / GC_TILE;
            total += pixel_loss(g, x, (uint32_t)y, t, 0xFFFFFFFFu, NULL, 0, &sig,
                                out_rgb ? out_rgb + 3 * ((size_t)y * g->W + x) : NULL);
        }
    if (loss) *loss = total;
}

/* One pixel for the difference quotients: the f32 forward (rasterize.ts:62-79
 * in f32, no contraction, on the f32 records, as the device evaluates it)
 * runs in lockstep and decides every branch: which entries composite, where
 * alpha saturates at the cap, where the loop stops and which output channels
 * clamp. The f64 values follow those decisions, giving the smooth function
 * whose derivative the device backward computes. *sig hashes what the f64
 * values alone would decide where they depend on the perturbed splat (its
 * own branch, the early stop, the output clamps): a difference between p+h
 * and p-h means a threshold lies inside the stencil. */
static double pixel_loss_forced(const orc_gradcheck *g, uint32_t x, uint32_t y, uint32_t t, uint32_t sub_rank,
                                const double *sub, double sub_radius, uint64_t *sig) {
    float T32 = 1.0f, C32[3] = {0, 0, 0};
    double T = 1.0, C[3] = {0, 0, 0};
    const float px32 = (float)x + 0.5f, py32 = (float)y + 0.5f;
    const double px = x + 0.5, py = y + 0.5;
    uint64_t h = 1469598103934665603ull;
    uint32_t step = 0;
    for (uint32_t e = g->toff[t]; e < g->toff[t + 1]; e++, step++) {
        if (T < kMinT_) h = (h ^ (uint64_t)(0x100000u + step)) * 1099511628211ull;
        if (T32 < (float)kMinT_) break;
        uint32_t k = g->tlist[e];
        const float *r = g->proj32 + 12 * (size_t)k;
        int code = 1;
        float a32 = 0.0f;
        {
            float dx = px32 - r[0], dy = py32 - r[1];
            if (!(r[2] <= 0.0f || fabsf(dx) > r[2] || fabsf(dy) > r[2])) {
                float power = orc_power32(r[4], r[5], r[6], dx, dy);
                if (power > 0.0f) code = 2;
                else {
                    float ge = orc_expf(power) - (float)kFloor;
                    float raw = r[7] * (ge > 0.0f ? ge : 0.0f);
                    a32 = raw < (float)kCap ? raw : (float)kCap;
                    code = a32 < (float)kMinA ? 3 : (raw < (float)kCap ? 4 : 5);
                }
            }
        }
        if (code < 4 && k != sub_rank) continue;
        /* other splats are constants of the checked function: their f32
         * records exactly as the forward computed them; only the perturbed
         * splat's record is recomputed (in f64) from its parameters */
        double other[9];
        if (k != sub_rank) {
            other[0] = r[0]; other[1] = r[1]; other[2] = r[4]; other[3] = r[5]; other[4] = r[6];
            other[5] = r[7]; other[6] = r[8]; other[7] = r[9]; other[8] = r[10];
        }
        const double *q = k == sub_rank ? sub : other;
        double raw64 = 0.0;
        int code64 = 1;
        {
            double rad = k == sub_rank ? sub_radius : g->radius[k];
            double dx = px - q[0], dy = py - q[1];
            if (!(rad <= 0 || fabs(dx) > rad || fabs(dy) > rad)) {
                double power = -0.5 * (q[2] * dx * dx + 2 * q[3] * dx * dy + q[4] * dy * dy);
                if (power > 0) code64 = 2;
                else {
                    double ge = exp(power) - kFloor;
                    raw64 = q[5] * (ge > 0 ? ge : 0);
                    double a64 = raw64 < kCap ? raw64 : kCap;
                    code64 = a64 < kMinA ? 3 : (raw64 < kCap ? 4 : 5);
                }
            }
        }
        if (k == sub_rank) h = (h ^ (uint64_t)(code64 + 8 * step)) * 1099511628211ull;
        if (code >= 4) {
            for (int c = 0; c < 3; c++) C32[c] = fmaf(T32 * a32, r[8 + c], C32[c]);
            T32 = T32 * (1.0f - a32);
            /* the f32 decision: saturated at the cap, or the unclamped value */
            double a = code == 5 ? kCap : raw64;
            for (int c = 0; c < 3; c++) C[c] += T * a * q[6 + c];
            T *= 1.0 - a;
        }
    }
    double loss = 0;
    for (int c = 0; c < 3; c++) {
        float o32 = C32[c] + T32 * (float)g->bg[c];
        double o = C[c] + T * g->bg[c];
        int in32 = o32 > 0.0f && o32 < 1.0f, in64 = o > 0 && o < 1;
        h = (h ^ (uint64_t)(16 + in64)) * 1099511628211ull;
        if (!in32) o = o32 <= 0.0f ? 0.0 : 1.0;
        double d = o - (double)g->target[3 * ((size_t)y * g->W + x) + c];
        loss += d * d;
    }
    *sig = h;
    return loss;
}

/* loss over the tiles of rank k with record rec for rank k; pixel
 * signatures combined in pixel order */
static double local_loss(const orc_gradcheck *g, uint32_t k, const double *rec, double rad, uint64_t *sig_out) {
    double total = 0;
    uint64_t sig = 0;
    for (uint32_t e = g->stoff[k]; e < g->stoff[k + 1]; e++) {
        uint32_t t = g->stiles[e];
        uint32_t tx = t % g->tiles_x, ty = t / g->tiles_x;
        for (uint32_t y = ty * GC_TILE; y < ty * GC_TILE + GC_TILE && y < g->H; y++)
            for (uint32_t x = tx * GC_TILE; x < tx * GC_TILE + GC_TILE && x < g->W; x++) {
                uint64_t ps;
                total += pixel_loss_forced(g, x, y, t, k, rec, rad, &ps);
                sig = (sig ^ ps) * 1099511628211ull + (x * 73856093u ^ y * 19349663u);
            }
    }
    *sig_out = sig;
    return total;
}

/* flags of orc_gc_fd */
#define GC_NOT_VISIBLE 1
#define GC_RADIUS 2
#define GC_ORDER 4
#define GC_CULL 8
#define GC_BRANCH 16
#define GC_PROJ_KINK 32

static int gc_fd_one(const orc_gradcheck *g, uint32_t id, uint32_t param, double h, double *out) {
    *out = 0;
    if (id >= g->s.count || param >= g->nparam) return -1;
    uint32_t k = g->rank[id];
    if (k == 0xFFFFFFFFu) return GC_NOT_VISIBLE;
    if (g->radius[k] <= 0) return GC_NOT_VISIBLE;
    double p[14 + 45], rec[2][9], rad[2], dep[2];
    uint32_t st[2] = {0, 0};
    uint64_t sg[2] = {0, 0};
    double L[2] = {0, 0};
    int flags = 0;
    for (int side = 0; side < 2; side++) {
        gc_params(g, id, p);
        p[param] += side ? h : -h;
        if (!proj64(g, p, rec[side], &rad[side], &dep[side], &st[side])) {
            flags |= GC_CULL;
            continue;
        }
        if (rad[side] != g->radius[k]) flags |= GC_RADIUS;
        /* the f32 forward's record moved by the exact f64 change of the
         * projection: the function differenced is the forward the device
         * evaluates (its f32 values) and its change is the f64 projection's */
        {
            const float *r32 = g->proj32 + 12 * (size_t)k;
            const double base32[9] = {r32[0], r32[1], r32[4], r32[5], r32[6], r32[7], r32[8], r32[9], r32[10]};
            const double *b64 = g->rec + 9 * (size_t)k;
            for (int f = 0; f < 9; f++) rec[side][f] = base32[f] + (rec[side][f] - b64[f]);
        }
        /* depth order: the f32 key of the f64 depth against the neighbours (ties by id) */
        if (param < 3) {
            uint32_t key = orc_float_to_sortable((float)dep[side]);
            if (k > 0) {
                uint32_t pk = g->keys[k - 1], pid = g->order[k - 1];
                if (key < pk || (key == pk && id < pid)) flags |= GC_ORDER;
            }
            if (k + 1 < g->nvis) {
                uint32_t nk = g->keys[k + 1], nid = g->order[k + 1];
                if (key > nk || (key == nk && id > nid)) flags |= GC_ORDER;
            }
        }
        L[side] = local_loss(g, k, rec[side], g->radius[k], &sg[side]);
    }
    if (flags & GC_CULL) return flags;
    if (sg[0] != sg[1]) flags |= GC_BRANCH;
    if (st[0] != st[1]) flags |= GC_PROJ_KINK;
    *out = (L[1] - L[0]) / (2 * h);
    return flags;
}

/* Difference quotient of L with respect to one field of the splat's f64
 * record (0 sx, 1 sy, 2 A, 3 B, 4 C, 5 alpha, 6-8 colour): the
 * screen-space gradients the backward's rasterizer pass produces. */
int orc_gc_fd_record(const orc_gradcheck *g, uint32_t id, uint32_t field, double h, double *out) {
    *out = 0;
    if (!g || id >= g->s.count || field >= 9) return -1;
    uint32_t k = g->rank[id];
    if (k == 0xFFFFFFFFu || g->radius[k] <= 0) return GC_NOT_VISIBLE;
    double rec[9], L[2];
    uint64_t sg[2];
    for (int side = 0; side < 2; side++) {
        memcpy(rec, g->rec + 9 * (size_t)k, sizeof(rec));
        rec[field] += side ? h : -h;
        L[side] = local_loss(g, k, rec, g->radius[k], &sg[side]);
    }
    *out = (L[1] - L[0]) / (2 * h);
    return sg[0] != sg[1] ? GC_BRANCH : 0;
}

/* dL/d(record alpha) of one splat by the backward's own recurrence (front
 * to back, S = what lies behind = output minus the running sum), in f64
 * (f64 = 1) or in f32 (f64 = 0), on the f32 forward's records and
 * decisions: separates the algorithm from its f32 evaluation. */
double orc_gc_dalpha(const orc_gradcheck *g, uint32_t id, int f64) {
    uint32_t k0 = g->rank[id];
    if (k0 == 0xFFFFFFFFu) return 0;
    double total = 0;
    for (uint32_t e = g->stoff[k0]; e < g->stoff[k0 + 1]; e++) {
        uint32_t t = g->stiles[e];
        uint32_t tx = t % g->tiles_x, ty = t / g->tiles_x;
        for (uint32_t y = ty * GC_TILE; y < ty * GC_TILE + GC_TILE && y < g->H; y++)
            for (uint32_t x = tx * GC_TILE; x < tx * GC_TILE + GC_TILE && x < g->W; x++) {
                /* forward in f32 (the device's values) */
                float T = 1.0f, Cc[3] = {0, 0, 0};
                const float px = (float)x + 0.5f, py = (float)y + 0.5f;
                uint32_t n = 0;
                for (uint32_t q = g->toff[t]; q < g->toff[t + 1]; q++, n++) {
                    if (T < (float)kMinT_) break;
                    const float *r = g->proj32 + 12 * (size_t)g->tlist[q];
                    float dx = px - r[0], dy = py - r[1];
                    if (r[2] <= 0.0f || fabsf(dx) > r[2] || fabsf(dy) > r[2]) continue;
                    float power = orc_power32(r[4], r[5], r[6], dx, dy);
                    if (power > 0.0f) continue;
                    float ge = orc_expf(power) - (float)kFloor;
                    float raw = r[7] * (ge > 0.0f ? ge : 0.0f);
                    float a = raw < (float)kCap ? raw : (float)kCap;
                    if (a < (float)kMinA) continue;
                    for (int c = 0; c < 3; c++) Cc[c] = fmaf(T * a, r[8 + c], Cc[c]);
                    T = T * (1.0f - a);
                }
                double G[3];
                for (int c = 0; c < 3; c++) {
                    float pre = Cc[c] + T * (float)g->bg[c];
                    float img = pre < 0.0f ? 0.0f : pre > 1.0f ? 1.0f : pre;
                    float grad = 2.0f * (img - g->target[3 * ((size_t)y * g->W + x) + c]);
                    G[c] = (pre > 0.0f && pre < 1.0f) ? grad : 0.0;
                }
                /* backward recurrence */
                double Td = 1.0, S[3];
                float Tf = 1.0f, Sf[3];
                for (int c = 0; c < 3; c++) { S[c] = Cc[c] + T * (float)g->bg[c]; Sf[c] = Cc[c] + T * (float)g->bg[c]; }
                for (uint32_t q = g->toff[t]; q < g->toff[t + 1]; q++) {
                    if (f64 ? Tf < (float)kMinT_ : Tf < (float)kMinT_) break;
                    uint32_t k = g->tlist[q];
                    const float *r = g->proj32 + 12 * (size_t)k;
                    float dx = px - r[0], dy = py - r[1];
                    if (r[2] <= 0.0f || fabsf(dx) > r[2] || fabsf(dy) > r[2]) continue;
                    float power = orc_power32(r[4], r[5], r[6], dx, dy);
                    if (power > 0.0f) continue;
                    float ef = orc_expf(power);
                    float ge = ef - (float)kFloor;
                    float raw = r[7] * (ge > 0.0f ? ge : 0.0f);
                    float a = raw < (float)kCap ? raw : (float)kCap;
                    if (a < (float)kMinA) continue;
                    if (f64) {
                        double w = Td * a, da = 0;
                        for (int c = 0; c < 3; c++) {
                            S[c] -= r[8 + c] * w;
                            da += G[c] * (r[8 + c] * Td - S[c] / (1.0 - a));
                        }
                        if (k == k0 && raw < (float)kCap) total += da * ((double)ef - kFloor);
                        Td *= 1.0 - a;
                    } else {
                        float w = Tf * a, da = 0.0f;
                        for (int c = 0; c < 3; c++) {
                            Sf[c] = Sf[c] - r[8 + c] * w;
                            da += (float)G[c] * (r[8 + c] * Tf - Sf[c] / (1.0f - a));
                        }
                        if (k == k0 && raw < (float)kCap) total += (double)(da * (ef - (float)kFloor));
                    }
                    Tf = Tf * (1.0f - a);
                }
            }
    }
    return total;
}

/* per pixel of one tile: the forced-loss quotient for the splat's alpha
 * (out_fd[256]) and the pixel's backward-recurrence term (out_an[256]) */
int orc_gc_pixel_terms(const orc_gradcheck *g, uint32_t id, uint32_t t, double h, double *out_fd, double *out_an) {
    uint32_t k0 = g->rank[id];
    if (k0 == 0xFFFFFFFFu) return -1;
    uint32_t tx = t % g->tiles_x, ty = t / g->tiles_x;
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t x = tx * GC_TILE + (i & 15), y = ty * GC_TILE + (i >> 4);
        out_fd[i] = out_an[i] = 0;
        if (x >= g->W || y >= g->H) continue;
        double rec[9];
        memcpy(rec, g->rec + 9 * (size_t)k0, sizeof(rec));
        uint64_t sg;
        double base5 = rec[5];
        rec[5] = base5 + h;
        double lp = pixel_loss_forced(g, x, y, t, k0, rec, g->radius[k0], &sg);
        rec[5] = base5 - h;
        double lm = pixel_loss_forced(g, x, y, t, k0, rec, g->radius[k0], &sg);
        out_fd[i] = (lp - lm) / (2 * h);
    }
    (void)out_an;
    return 0;
}

int orc_gc_fd(const orc_gradcheck *g, uint32_t n, const uint32_t *ids, const uint32_t *params, double h,
              double *out, int32_t *flags, int nthreads) {
    if (!g) return -1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; i++) flags[i] = gc_fd_one(g, ids[i], params[i], h, out + i);
    return 0;
}
