// Device code shared by the single-view renderer (render.hip) and the
// multi-view batch (batch.hip): projection, cull, the rasterizer's tile loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "dl_common.h"
#include "render.h"

namespace dl {
namespace {

// ---------------------------------------------------------------------------
// project (project.ts:22-332)
// ---------------------------------------------------------------------------
__device__ inline void quat_rot(const float q[4], float v0, float v1, float v2, float o[3]) {
  const float c0 = q[2] * v2 - q[3] * v1;
  const float c1 = q[3] * v0 - q[1] * v2;
  const float c2 = q[1] * v1 - q[2] * v0;
  const float t0 = 2.0f * c0, t1 = 2.0f * c1, t2 = 2.0f * c2;
  const float d0 = q[2] * t2 - q[3] * t1;
  const float d1 = q[3] * t0 - q[1] * t2;
  const float d2 = q[1] * t1 - q[2] * t0;
  o[0] = (v0 + q[0] * t0) + d0;
  o[1] = (v1 + q[0] * t1) + d1;
  o[2] = (v2 + q[0] * t2) + d2;
}

template <int BANDS>
__device__ inline void eval_sh(const float *__restrict__ sh, float x, float y, float z, float out[3]) {
  constexpr int M = BANDS == 1 ? 3 : BANDS == 2 ? 8 : 15;
  float c[3 * M];
#pragma unroll
  for (int k = 0; k < 3 * M; k++) c[k] = sh[k];
  float r = 0.0f, g = 0.0f, b = 0.0f;
  const float SH1 = (float)0.4886025119029199;
  const float e0 = (-SH1) * y, e1 = SH1 * z, e2 = (-SH1) * x;
  r = r + e0 * c[0]; r = r + e1 * c[1]; r = r + e2 * c[2];
  g = g + e0 * c[M + 0]; g = g + e1 * c[M + 1]; g = g + e2 * c[M + 2];
  b = b + e0 * c[2 * M + 0]; b = b + e1 * c[2 * M + 1]; b = b + e2 * c[2 * M + 2];
  if constexpr (BANDS >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    float b2[5];
    b2[0] = (float)1.0925484305920792 * xy;
    b2[1] = (float)-1.0925484305920792 * yz;
    b2[2] = (float)0.31539156525252005 * ((2.0f * zz - xx) - yy);
    b2[3] = (float)-1.0925484305920792 * xz;
    b2[4] = (float)0.5462742152960396 * (xx - yy);
#pragma unroll
    for (int k = 0; k < 5; k++) {
      r = r + b2[k] * c[3 + k];
      g = g + b2[k] * c[M + 3 + k];
      b = b + b2[k] * c[2 * M + 3 + k];
    }
    if constexpr (BANDS >= 3) {
      float b3[7];
      b3[0] = ((float)-0.5900435899266435 * y) * (3.0f * xx - yy);
      b3[1] = ((float)2.890611442640554 * xy) * z;
      b3[2] = ((float)-0.4570457994644658 * y) * ((4.0f * zz - xx) - yy);
      b3[3] = ((float)0.3731763325901154 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
      b3[4] = ((float)-0.4570457994644658 * x) * ((4.0f * zz - xx) - yy);
      b3[5] = ((float)1.445305721320277 * z) * (xx - yy);
      b3[6] = ((float)-0.5900435899266435 * x) * (xx - 3.0f * yy);
#pragma unroll
      for (int k = 0; k < 7; k++) {
        r = r + b3[k] * c[8 + k];
        g = g + b3[k] * c[M + 8 + k];
        b = b + b3[k] * c[2 * M + 8 + k];
      }
    }
  }
  out[0] = r;
  out[1] = g;
  out[2] = b;
}

// Camera-space position of a splat centre (project.ts:80-82).
__device__ inline void cam_pos(const float *m, float px, float py, float pz, float &pcx, float &pcy, float &pcz) {
  pcx = ((m[0] * px + m[4] * py) + m[8] * pz) + m[12] * 1.0f;
  pcy = ((m[1] * px + m[5] * py) + m[9] * pz) + m[13] * 1.0f;
  pcz = ((m[2] * px + m[6] * py) + m[10] * pz) + m[14] * 1.0f;
}

// Colour of a splat (project.ts:203-212): DC + view-dependent SH along the
// camera-space direction, + 0.5, clamped at 0.
template <int BANDS>
__device__ inline void color_from(const float *dc, const float *shc, float pcx, float pcy, float pcz, float col[3]) {
  col[0] = dc[0];
  col[1] = dc[1];
  col[2] = dc[2];
  if constexpr (BANDS > 0) {
    const float len = __builtin_sqrtf((pcx * pcx + pcy * pcy) + pcz * pcz);
    float sh[3];
    eval_sh<BANDS>(shc, pcx / len, pcy / len, pcz / len, sh);
    col[0] = col[0] + sh[0];
    col[1] = col[1] + sh[1];
    col[2] = col[2] + sh[2];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) col[k] = wmax(col[k] + 0.5f, 0.0f);
}

template <int BANDS>
__device__ inline void splat_color(const SceneView &sc, uint32_t id, float pcx, float pcy, float pcz, float col[3]) {
  const float *r = static_cast<const float *>(
      __builtin_assume_aligned(sc.col + (size_t)id * color_record_floats(BANDS), 16));
  color_from<BANDS>(r, r + 3, pcx, pcy, pcz, col);
}

// One splat through project.ts:59-234. Returns the tile count; a culled
// splat gets slot 0 = (0, 0, -1, 0) and count 0 (project.ts:165-169, 196-200, 238-242, 253-257).
// Without COLOR the colour lanes of slot 2 stay 0: k_splat_colors fills them
// later for the splats that reach a tile list.
template <int BANDS, bool COLOR = true>
__device__ inline uint32_t project_splat(const SceneView &sc, uint32_t id, const ProjUniforms &u, float4 &o0,
                                         float4 &o1, float4 &o2, bool &cov16) {
  cov16 = false;
  const float *m = u.view;
  o0 = make_float4(0.0f, 0.0f, -1.0f, 0.0f);
  o1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  o2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float px = sc.pos[3 * (size_t)id], py = sc.pos[3 * (size_t)id + 1], pz = sc.pos[3 * (size_t)id + 2];
  const float4 qv = reinterpret_cast<const float4 *>(sc.rot)[id];
  const float q[4] = {qv.x, qv.y, qv.z, qv.w};
  const float s0 = dl_expf(sc.ls[3 * (size_t)id]);
  const float s1 = dl_expf(sc.ls[3 * (size_t)id + 1]);
  const float s2 = dl_expf(sc.ls[3 * (size_t)id + 2]);
  const float logit = sc.op[id];

  float pcx, pcy, pcz;
  cam_pos(m, px, py, pz, pcx, pcy, pcz);
  const float depth = -pcz;
  float sx, sy;
  if (u.proj == DL_PINHOLE) {
    if (depth <= u.near_ || depth >= u.far_) return 0;
    const float proj_x = (pcx * u.focal) / (-pcz);
    const float proj_y = ((-pcy) * u.focal) / (-pcz);
    sx = proj_x + u.cx;
    sy = proj_y + u.cy;
  } else {
    const float rr = __builtin_sqrtf((pcx * pcx + pcy * pcy) + pcz * pcz);
    if (rr <= u.near_) return 0;
    const float lon = (float)atan2((double)pcx, (double)(-pcz));
    const float lat = (float)asin((double)(pcy / rr));
    sx = (lon / kTwoPi + 0.5f) * u.width;
    sy = (0.5f - lat / kPi) * u.height;
  }
  float rx[3], ry[3], rz[3];
  quat_rot(q, 1.0f, 0.0f, 0.0f, rx);
  quat_rot(q, 0.0f, 1.0f, 0.0f, ry);
  quat_rot(q, 0.0f, 0.0f, 1.0f, rz);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    rx[k] = rx[k] * s0;
    ry[k] = ry[k] * s1;
    rz[k] = rz[k] * s2;
  }
  const float c00 = (rx[0] * rx[0] + ry[0] * ry[0]) + rz[0] * rz[0];
  const float c01 = (rx[0] * rx[1] + ry[0] * ry[1]) + rz[0] * rz[1];
  const float c02 = (rx[0] * rx[2] + ry[0] * ry[2]) + rz[0] * rz[2];
  const float c11 = (rx[1] * rx[1] + ry[1] * ry[1]) + rz[1] * rz[1];
  const float c12 = (rx[1] * rx[2] + ry[1] * ry[2]) + rz[1] * rz[2];
  const float c22 = (rx[2] * rx[2] + ry[2] * ry[2]) + rz[2] * rz[2];
  const float v00 = m[0], v01 = m[4], v02 = m[8];
  const float v10 = m[1], v11 = m[5], v12 = m[9];
  const float v20 = m[2], v21 = m[6], v22 = m[10];
  const float a00 = (v00 * c00 + v01 * c01) + v02 * c02;
  const float a01 = (v00 * c01 + v01 * c11) + v02 * c12;
  const float a02 = (v00 * c02 + v01 * c12) + v02 * c22;
  const float a10 = (v10 * c00 + v11 * c01) + v12 * c02;
  const float a11 = (v10 * c01 + v11 * c11) + v12 * c12;
  const float a12 = (v10 * c02 + v11 * c12) + v12 * c22;
  const float a20 = (v20 * c00 + v21 * c01) + v22 * c02;
  const float a21 = (v20 * c01 + v21 * c11) + v22 * c12;
  const float a22 = (v20 * c02 + v21 * c12) + v22 * c22;
  const float cc00 = (a00 * v00 + a01 * v01) + a02 * v02;
  const float cc01 = (a00 * v10 + a01 * v11) + a02 * v12;
  const float cc02 = (a00 * v20 + a01 * v21) + a02 * v22;
  const float cc11 = (a10 * v10 + a11 * v11) + a12 * v12;
  const float cc12 = (a10 * v20 + a11 * v21) + a12 * v22;
  const float cc22 = (a20 * v20 + a21 * v21) + a22 * v22;
  const float inv_z = 1.0f / pcz;
  const float j00 = u.focal * inv_z;
  const float j02 = (((-u.focal) * pcx) * inv_z) * inv_z;
  const float j11 = u.focal * inv_z;
  const float j12 = (((-u.focal) * pcy) * inv_z) * inv_z;
  const float b00 = j00 * cc00 + j02 * cc02;
  const float b01 = j00 * cc01 + j02 * cc12;
  const float b02 = j00 * cc02 + j02 * cc22;
  const float b11 = j11 * cc11 + j12 * cc12;
  const float b12 = j11 * cc12 + j12 * cc22;
  float cov00 = b00 * j00 + b02 * j02;
  const float cov01 = b01 * j11 + b02 * j12;
  float cov11 = b11 * j11 + b12 * j12;
  const float det_orig = cov00 * cov11 - cov01 * cov01;
  cov00 = cov00 + kAA;
  cov11 = cov11 + kAA;
  const float det = cov00 * cov11 - cov01 * cov01;
  const float aa = __builtin_sqrtf(wmax(det_orig / det, 0.0f));
  if (det <= kDetMin) return 0;
  const float inv_det = 1.0f / det;
  const float conic_a = cov11 * inv_det;
  const float conic_b = (-cov01) * inv_det;
  const float conic_c = cov00 * inv_det;
  const float mid = 0.5f * (cov00 + cov11);
  const float r_eig = __builtin_sqrtf(wmax(kEigMin, mid * mid - det));
  const float lambda1 = mid + r_eig;
  const float radius_raw = kSigma * __builtin_sqrtf(lambda1);
  const float fade_s = kFadeStart * u.height;
  const float fade_e = kFadeEnd * u.height;
  const float fade = wclamp((fade_e - radius_raw) / (fade_e - fade_s), 0.0f, 1.0f);
  if (fade <= 0.0f) return 0;
  const float radius = __builtin_ceilf(wmin(radius_raw, fade_e));
  // without COLOR the colour lanes carry the camera-space position for
  // k_splat_colors
  float col[3] = {pcx, pcy, pcz};
  if constexpr (COLOR) splat_color<BANDS>(sc, id, pcx, pcy, pcz, col);
  const float alpha = (dl_sigmoidf(logit) * fade) * aa;
  // Rasterizer skip threshold in the record's padding lane (0 in the
  // reference, project.ts:221): for power < tau the reference's pixel alpha,
  // min(cap, alpha*max(0, exp(power) - floor)), is below 1/255. tau is
  // ln(floor + (1/255)/alpha) lowered by margins that cover exp's error and
  // the f32 roundings of that expression.
  float tau = __builtin_inff();
  if (alpha > 0.0f) {
    const double target = (double)kGaussFloor + ((double)kMinAlpha / (double)alpha) * (1.0 - 1e-6);
    const double t = log(target) - 1e-5;
    tau = (float)(t - fabs(t) * 1e-6 - 1e-30);
  }
  o0 = make_float4(sx, sy, radius, depth);
  o1 = make_float4(conic_a, conic_b, conic_c, alpha);
  o2 = make_float4(col[0], col[1], col[2], tau);
  // tile coverage (project.ts:223-233), at the grid's tile size
  uint32_t min_tx, max_tx, min_ty, max_ty, count = 0;
  if (tile_range(sx, sy, radius, u.tg, min_tx, max_tx, min_ty, max_ty)) {
    count = (max_tx - min_tx + 1u) * (max_ty - min_ty + 1u);
    if (count > u.max_cov) count = u.max_cov;
  }
  // would the reference (16-px tiles) truncate this splat's pairs?
  tile_range16(sx, sy, radius, u.tg.t16x, u.tg.t16y, min_tx, max_tx, min_ty, max_ty);
  cov16 = max_tx >= min_tx && max_ty >= min_ty &&
          (uint64_t)(max_tx - min_tx + 1u) * (max_ty - min_ty + 1u) > u.max_cov16;
  return count;
}


// f64 cull (camera.ts:56-99) and sortable depth key of splat i, then
// project() when it is visible; the outputs go to slot o of keys / vals /
// rec / counts (counts may be null), with vals[o] = val.
template <int BANDS, bool COLOR = true>
__device__ inline void cull_project_one(const SceneView &sc, const ProjUniforms &u, double near_plane,
                                        double far_plane, uint32_t i, size_t o, uint32_t val,
                                        uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                        float4 *__restrict__ rec, uint32_t *__restrict__ counts, bool &vis,
                                        uint32_t &count, bool &cov16) {
  const float *m = u.view;
  const double x = sc.pos[3 * (size_t)i], y = sc.pos[3 * (size_t)i + 1], z = sc.pos[3 * (size_t)i + 2];
  float depth_f;
  if (u.proj == DL_PINHOLE) {
    double cz = (double)m[2] * x + (double)m[6] * y;
    cz = cz + (double)m[10] * z;
    cz = cz + (double)m[14];
    const double depth = -cz;
    vis = depth > near_plane && depth < far_plane;
    depth_f = (float)depth;
  } else {
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
    vis = r2 > near_plane * near_plane;
    depth_f = (float)__builtin_sqrt(r2);
  }
  keys[o] = vis ? sortable_key(depth_f) : kInvisibleKey;
  if (vals) vals[o] = val;
  float4 o0 = make_float4(0.0f, 0.0f, -1.0f, 0.0f), o1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), o2 = o1;
  count = 0;
  cov16 = false;
  if (vis) count = project_splat<BANDS, COLOR>(sc, i, u, o0, o1, o2, cov16);
  rec[3 * o + 0] = o0;
  rec[3 * o + 1] = o1;
  rec[3 * o + 2] = o2;
  if (counts) counts[o] = count;
}

// Colours of the splats of one pass, computed after a projection without
// COLOR with the same arithmetic, from the camera-space position that the
// projection left in the colour lanes: item j < n names splat ids[j] (or
// rects[j].x), n = the static n_cap or min(n_cap, *d_n). The colours of
// culled splats (in no tile list) are meaningless and never read.
template <int BANDS>
__global__ __launch_bounds__(256) void k_splat_colors(SceneView sc, const uint32_t *__restrict__ ids,
                                                      const uint4 *__restrict__ rects, uint32_t n_cap,
                                                      const uint32_t *__restrict__ d_n, float4 *__restrict__ rec) {
  const uint32_t n = d_n ? min(*d_n, n_cap) : n_cap;
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const uint32_t id = ids ? ids[j] : rects[j].x;
  float *slot = reinterpret_cast<float *>(rec + 3 * (size_t)id + 2);
  const float pcx = slot[0], pcy = slot[1], pcz = slot[2];
  float col[3];
  splat_color<BANDS>(sc, id, pcx, pcy, pcz, col);
  slot[0] = col[0];
  slot[1] = col[1];
  slot[2] = col[2];
}

// ---------------------------------------------------------------------------
// rasterize (rasterize.ts:36-82) + finalize (finalize.ts:29-42)
//
// One wave64 per 16x16 tile. Lane l owns pixel column l%16 and the four rows
// 4*(l/16) .. 4*(l/16)+3, so the wave covers the whole tile and needs no
// workgroup barrier. The tile's splats are staged 32 at a time in LDS and read
// back as broadcasts; the strips each staged splat can reach are four 32-bit
// ballots held in scalar registers, and the wave visits only the splats with
// a bit set. Every pixel runs exactly the reference's loop body in the
// reference's order; a pixel stops when T < 1e-4 and the wave stops when all
// its pixels have.
// ---------------------------------------------------------------------------
__device__ inline uint32_t pack_rgba(float cr, float cg, float cb, float T, const RastParams &p) {
  const float r = wclamp(cr + T * p.bg_r, 0.0f, 1.0f);
  const float g = wclamp(cg + T * p.bg_g, 0.0f, 1.0f);
  const float b = wclamp(cb + T * p.bg_b, 0.0f, 1.0f);
  const float a = wclamp((1.0f - T) + T * p.bg_a, 0.0f, 1.0f);
  const uint32_t R = u32sat(r * 255.0f + 0.5f);
  const uint32_t G = u32sat(g * 255.0f + 0.5f);
  const uint32_t B = u32sat(b * 255.0f + 0.5f);
  const uint32_t A = u32sat(a * 255.0f + 0.5f);
  return R | (G << 8) | (B << 16) | (A << 24);
}

// The Gaussian exponent of rasterize.ts:71, -0.5 (A dx^2 + 2B dx dy + C dy^2),
// with the two contractions the specification fixes (WGSL lets a compiler
// fuse a * b + c): fma(C dy, dy, fma(2B dx, dy, A dx dx)). The colour
// accumulation is the other fused expression. Both the backward's forward
// recomputation and the oracle use exactly this form.
__device__ inline float gauss_power(float A, float B, float C, float dx, float dy) {
  return (-0.5f) * __builtin_fmaf(C * dy, dy, __builtin_fmaf((2.0f * B) * dx, dy, (A * dx) * dx));
}

// Ascending 6-bit sort key of a tile's work: 63 for none, then two bands per
// octave, heavier work first.
__device__ inline uint32_t work_key(uint32_t w) {
  if (w == 0) return 63u;
  const uint32_t lz = 31u - (uint32_t)__clz(w);
  const uint32_t half = lz >= 1 ? (w >> (lz - 1)) & 1u : 0u;
  return 62u - min(62u, 2u * lz + half);
}

// LDS of one rasterizer wave: a batch of kStage staged splat records, the
// lanes' 16 accumulators while the batch is staged, and a window of splat
// ids (32-px tiles).
constexpr uint32_t kStage = 32;
struct RastLds {
  float4 rec[kStage * 3];
  float acc[16 * 64];
  uint32_t ids[64];  // the 32-px rasterizer's filtered window
};

// Synchronisation of a staging batch. A one-wave workgroup takes the
// workgroup barrier. In a workgroup of several waves each wave stages its own
// batches and leaves its loop on its own pixels, so only the wave's own lanes
// must see the staged records: the same fences without the s_barrier.
template <bool MULTI>
__device__ inline void stage_sync() {
  if constexpr (MULTI) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// A tile that a pass skips: closed (and final, or not the last pass), or
// without entries in a non-final pass while the accumulator holds its state.
__device__ inline bool tile_skipped(const RastParams &p, uint32_t vtile, uint32_t start, uint32_t end) {
  return p.closed && p.resume && ((p.closed[vtile] && (p.final_on_close || !p.finalize)) || (!p.finalize && start == end));
}

// The pixels of one wave over a 16x16 region at (ox, oy): lane l owns column
// l % 16 and rows l / 16 + 4k (strip k = 0..3). Pixels outside the image
// start with T = 0, so they are never live and never written.
struct Px16 {
  float cr[4], cg[4], cb[4], T[4], pyf[4];
  float pxf;
  uint32_t px, py0;
  bool inb[4];
};

__device__ inline void px16_load(Px16 &s, uint32_t ox, uint32_t oy, const float4 *__restrict__ accum,
                                 const RastParams &p) {
  const int lane = threadIdx.x & 63;
  s.px = ox + (lane & 15);
  s.py0 = oy + (lane >> 4);
  s.pxf = (float)s.px + 0.5f;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t y = s.py0 + 4 * k;
    s.inb[k] = s.px < p.width && y < p.height;
    s.pyf[k] = (float)y + 0.5f;
    s.cr[k] = 0.0f; s.cg[k] = 0.0f; s.cb[k] = 0.0f; s.T[k] = s.inb[k] ? 1.0f : 0.0f;
    if (s.inb[k] && p.resume) {
      const float4 a = accum[(size_t)y * p.width + s.px];
      s.cr[k] = a.x; s.cg[k] = a.y; s.cb[k] = a.z; s.T[k] = a.w;
    }
  }
}

__device__ inline bool px16_open(const Px16 &s) {
  return __ballot(!(s.T[0] < kMinT) || !(s.T[1] < kMinT) || !(s.T[2] < kMinT) || !(s.T[3] < kMinT)) != 0;
}

__device__ inline void px16_store(const Px16 &s, const RastParams &p, float4 *__restrict__ accum,
                                  uint32_t *__restrict__ out_rgba, bool write_accum, bool finalize) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!s.inb[k]) continue;
    const size_t pix = (size_t)(s.py0 + 4 * k) * p.width + s.px;
    if (write_accum) accum[pix] = make_float4(s.cr[k], s.cg[k], s.cb[k], s.T[k]);
    if (finalize) out_rgba[pix] = pack_rgba(s.cr[k], s.cg[k], s.cb[k], s.T[k], p);
  }
}

// The reference's per-pixel loop body (rasterize.ts:63-78) for one splat,
// split at the pixel column: a lane's pixels share one column, so the box
// test on dx and the dx-only products of the exponent, (2B) dx and
// (A dx) dx, are computed once per splat (col_terms) for all of the lane's
// rows. blend_row then evaluates gauss_power's expression with those
// products, the same operations in the same order.
struct ColTerms {
  float bdx, adx2;
  bool out;  // r <= 0 or |dx| > r: every row of the column skips the splat
};

__device__ inline ColTerms col_terms(const float4 a, const float4 b, float pxf) {
  const float dx = pxf - a.x;
  const float r = a.z;
  ColTerms c;
  c.out = r <= 0.0f || __builtin_fabsf(dx) > r;
  c.bdx = (2.0f * b.y) * dx;
  c.adx2 = (b.x * dx) * dx;
  return c;
}

__device__ inline void blend_row(const ColTerms &c, const float4 a, const float4 b, const float4 col, float pyf,
                                 float &cr, float &cg, float &cb, float &T) {
  // The pixel's own tests: saturated (rasterize.ts:63), then the box test
  // (:69). One branch on them; the exponent, alpha and their tests run
  // without further branches and the blend is taken where all pass (nested
  // branches per test cost 5% more on the 6M-splat 4K frame).
  const float dy = pyf - a.y;
  if (T < kMinT || c.out || __builtin_fabsf(dy) > a.z) return;
  const float power = (-0.5f) * __builtin_fmaf(b.z * dy, dy, __builtin_fmaf(c.bdx, dy, c.adx2));
  // min(cap, opacity * max(0, exp(power) - floor)) (rasterize.ts:73)
  // without the max: the record's opacity is sigmoid * fade * aa, so
  // >= 0 or NaN, and for exp - floor <= 0 the product is <= 0 and
  // fails the 1/255 test below exactly like 0 does (NaN stays NaN). The exp
  // of a power outside [col.w, 0] is computed and not used.
  const float alpha = wmin(kOpacityCap, b.w * (dl_expf_nonpos(power) - kGaussFloor));
  if (!(power > 0.0f) && !(power < col.w) && !(alpha < kMinAlpha)) {
    const float w = T * alpha;
    cr = __builtin_fmaf(w, col.x, cr);
    cg = __builtin_fmaf(w, col.y, cg);
    cb = __builtin_fmaf(w, col.z, cb);
    T = T * (1.0f - alpha);
  }
}

// Strips (0..3) of a wave's 16x16 region that still hold a live pixel.
__device__ inline uint32_t px16_live(const Px16 &s) {
  uint32_t live = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) live |= (__ballot(!(s.T[k] < kMinT)) != 0 ? 1u : 0u) << k;
  return live;
}

// One staged batch for a wave's 16x16 region with rows from oy: slots
// c < cnt (cnt <= 32) hold splat get_j(c), in list order. The records are
// staged in LDS and read back as broadcasts; the strips each staged splat
// can reach are four 32-bit ballots, and the wave visits only the splats
// with a bit set (COUNT: every staged splat, counting the reference's loop
// iterations of the live pixels). Every pixel runs exactly the reference's
// loop body in the reference's order. Returns 8 plus the visited (splat,
// strip) bits.
template <bool COUNT, bool MULTI, class GetJ>
__device__ inline uint32_t px16_batch(Px16 &s, uint32_t oy, float tile_x0, float tile_x1, uint32_t cnt,
                                      uint32_t strips_live, GetJ get_j, const float4 *__restrict__ rec, RastLds &L,
                                      uint32_t &iters) {
  const int lane = threadIdx.x & 63;
  // Staging: the accumulators wait in LDS while the strip tests run,
  // so they hold no VGPRs there (78 VGPRs instead of 94); the batch is 32
  // splats, two lanes per splat with two strips each, so that the LDS
  // of a wave stays under 6 KB. Both bounds together allow 6 waves per
  // SIMD.
#pragma unroll
  for (int k = 0; k < 4; k++) {
    L.acc[(4 * k + 0) * 64 + lane] = s.cr[k];
    L.acc[(4 * k + 1) * 64 + lane] = s.cg[k];
    L.acc[(4 * k + 2) * 64 + lane] = s.cb[k];
    L.acc[(4 * k + 3) * 64 + lane] = s.T[k];
  }
  const uint32_t slot = (uint32_t)lane & (kStage - 1u), half = (uint32_t)lane / kStage;
  uint32_t m = 0;
  if (slot < cnt) {
    const uint32_t j = get_j(slot);
    const float4 a = rec[3 * (size_t)j + 0];
    const float4 b = rec[3 * (size_t)j + 1];
    const float4 col = rec[3 * (size_t)j + 2];
    if (half == 0) {
      L.rec[3 * slot + 0] = a;
      L.rec[3 * slot + 1] = b;
      L.rec[3 * slot + 2] = col;
    }
    const StripTest st = strip_test(a.x, a.y, a.z, b.x, b.y, b.z, col.w, tile_x0, tile_x1);
#pragma unroll
    for (uint32_t q = 0; q < 2; q++) {
      const uint32_t k = 2 * half + q;
      const float y0 = (float)(oy + 4 * k) + 0.5f;
      if (strip_may_hit(st, y0, y0 + 3.0f)) m |= 1u << k;
    }
  }
  // bit c of sk: staged splat c may reach strip k (lanes 0-31 test
  // strips 0 and 1, lanes 32-63 strips 2 and 3)
  const uint64_t b02 = __ballot((m & 0x5u) != 0u), b13 = __ballot((m & 0xAu) != 0u);
  const uint32_t s0 = (strips_live & 1u) ? (uint32_t)b02 : 0u;
  const uint32_t s1 = (strips_live & 2u) ? (uint32_t)b13 : 0u;
  const uint32_t s2 = (strips_live & 4u) ? (uint32_t)(b02 >> 32) : 0u;
  const uint32_t s3 = (strips_live & 8u) ? (uint32_t)(b13 >> 32) : 0u;
  const uint32_t work = 8u + (uint32_t)(__builtin_popcount(s0) + __builtin_popcount(s1) +
                                        __builtin_popcount(s2) + __builtin_popcount(s3));
  stage_sync<MULTI>();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    s.cr[k] = L.acc[(4 * k + 0) * 64 + lane];
    s.cg[k] = L.acc[(4 * k + 1) * 64 + lane];
    s.cb[k] = L.acc[(4 * k + 2) * 64 + lane];
    s.T[k] = L.acc[(4 * k + 3) * 64 + lane];
  }
  // lane c (and c + 32) holds staged splat c's four strip bits; the loop
  // reads them with one readlane instead of assembling them from s0..s3 in
  // scalar registers (~14 scalar instructions per visited splat; the
  // rasterize stage took 5% longer that way)
  const uint32_t lc = (uint32_t)lane & 31u;
  const uint32_t mvec = ((s0 >> lc) & 1u) | (((s1 >> lc) & 1u) << 1) | (((s2 >> lc) & 1u) << 2) |
                        (((s3 >> lc) & 1u) << 3);
  uint32_t todo = COUNT ? (cnt == 32u ? ~0u : (1u << cnt) - 1u) : (s0 | s1 | s2 | s3);
  while (todo) {
    const uint32_t c = (uint32_t)__builtin_ctz(todo);
    todo &= todo - 1u;
    if (COUNT)
      iters += (uint32_t)!(s.T[0] < kMinT) + (uint32_t)!(s.T[1] < kMinT) + (uint32_t)!(s.T[2] < kMinT) +
               (uint32_t)!(s.T[3] < kMinT);
    const uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)mvec, (int)c);
    if (mm == 0) continue;
    const float4 a = L.rec[3 * c + 0];
    const float4 b = L.rec[3 * c + 1];
    const float4 col = L.rec[3 * c + 2];
    // the strip bits are uniform over the wave (scalar branches); the
    // column test and the saturation test are per lane
    const ColTerms ct = col_terms(a, b, s.pxf);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (!((mm >> k) & 1u)) continue;
      blend_row(ct, a, b, col, s.pyf[k], s.cr[k], s.cg[k], s.cb[k], s.T[k]);
    }
  }
  stage_sync<MULTI>();
  return work;
}

// Front to back over list entries [start, end) of pvals (which index rec)
// for one wave's 16x16 region at (ox, oy), 32 entries per batch; a pixel
// stops when T < 1e-4 and the wave stops when all its pixels have. Returns
// the sum of the batches' work; consumed counts the entries read before
// the stop.
template <bool COUNT, bool MULTI>
__device__ inline uint32_t px16_run(Px16 &s, uint32_t ox, uint32_t oy, uint32_t start, uint32_t end,
                                    const float4 *__restrict__ rec, const uint32_t *__restrict__ pvals, RastLds &L,
                                    uint32_t &iters, uint32_t &consumed) {
  uint32_t work = 0;
  consumed = 0;
  if (start >= end) return 0;
  const float tile_x0 = (float)ox + 0.5f, tile_x1 = tile_x0 + 15.0f;
  uint32_t stop = end;
  for (uint32_t base = start; base < end; base += kStage) {
    const uint32_t strips_live = px16_live(s);
    if (strips_live == 0) {
      stop = base;
      break;
    }
    work += px16_batch<COUNT, MULTI>(s, oy, tile_x0, tile_x1, min(kStage, end - base), strips_live,
                                     [&](uint32_t c) { return pvals[base + c]; }, rec, L, iters);
  }
  consumed = stop - start;
  return work;
}

// The 16-px quadrant (q16x, q16y) of a 32-px tile whose list is entries
// [start, end) of pvals: windows of 64 entries are read for the splats'
// centres and radii, the entries whose 16-px range holds the quadrant (the
// reference's list of that 16-px tile) are kept in list order by ballot,
// and only those are staged, 32 per batch. Returns the batches' work.
template <bool MULTI>
__device__ inline uint32_t px16_run_quad(Px16 &s, uint32_t ox, uint32_t oy, uint32_t start, uint32_t end,
                                         const float4 *__restrict__ rec, const uint32_t *__restrict__ pvals,
                                         const RastParams &p, RastLds &L, uint32_t q16x, uint32_t q16y) {
  uint32_t work = 0, iters = 0;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  const float tile_x0 = (float)ox + 0.5f, tile_x1 = tile_x0 + 15.0f;
  for (uint32_t base = start; base < end; base += 64u) {
    uint32_t strips_live = px16_live(s);
    if (strips_live == 0) break;
    const uint32_t i = base + (uint32_t)lane;
    bool listed = false;
    uint32_t j = 0;
    if (i < end) {
      j = pvals[i];
      const float4 a = rec[3 * (size_t)j];
      uint32_t x0, x1, y0, y1;
      tile_range16(a.x, a.y, a.z, p.t16x, p.t16y, x0, x1, y0, y1);
      listed = q16x >= x0 && q16x <= x1 && q16y >= y0 && q16y <= y1;
    }
    const uint64_t lm = __ballot(listed);
    if (listed) L.ids[__popcll(lm & lt)] = j;
    stage_sync<MULTI>();
    const uint32_t nl = (uint32_t)__popcll(lm);
    for (uint32_t c0 = 0; c0 < nl; c0 += kStage) {
      if (c0 > 0) {
        strips_live = px16_live(s);
        if (strips_live == 0) break;
      }
      work += px16_batch<false, MULTI>(s, oy, tile_x0, tile_x1, min(kStage, nl - c0), strips_live,
                                       [&](uint32_t c) { return L.ids[c0 + c]; }, rec, L, iters);
    }
  }
  return work;
}

// The work of one wave64 on one 16-px tile: tile (tx, ty) of the view, whose
// accum / out_rgba are the view's own row-major pixels, with list entries
// [start, end) of pvals (which index rec); vtile
// indexes the open/closed flags (p.closed). A tile is closed once all its
// pixels have T < 1e-4: later passes give it no entries, so its colour is
// final. With p.final_on_close it writes its RGBA8 when it closes (and no
// accumulator) and every later pass skips it; otherwise later passes skip it
// until the final one, which finalizes it from the accumulator. A tile
// without entries in a non-final pass is skipped too, as long as the
// accumulator already holds its state. Returns the tile's work measure
// (0 when skipped).
template <bool COUNT, bool MULTI>
__device__ inline uint32_t rasterize_tile16(uint32_t tx, uint32_t ty, uint32_t vtile, uint32_t start, uint32_t end,
                                            const float4 *__restrict__ rec, const uint32_t *__restrict__ pvals,
                                            float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba,
                                            const RastParams &p, unsigned long long *__restrict__ d_iters,
                                            RastLds &L) {
  if (tile_skipped(p, vtile, start, end)) return 0;
  const int lane = threadIdx.x & 63;
  Px16 s;
  px16_load(s, tx * kTile, ty * kTile, accum, p);
  uint32_t iters = 0, consumed = 0;
  const uint32_t work = px16_run<COUNT, MULTI>(s, tx * kTile, ty * kTile, start, end, rec, pvals, L, iters, consumed);
  bool closing = false;
  if (p.closed && !p.finalize) {
    const bool open = px16_open(s);
    if (lane == 0) p.closed[vtile] = open ? 0 : 1;
    closing = !open && p.final_on_close;
  }
  px16_store(s, p, accum, out_rgba, p.write_accum && !closing, p.finalize || closing);
  if (COUNT) {
    uint32_t sum = iters;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) {
      atomicAdd(d_iters, (unsigned long long)sum);
      atomicAdd(d_iters + 1, (unsigned long long)consumed);
      if (p.row_iters) atomicAdd(p.row_iters + ty, (unsigned long long)sum);
    }
  }
  return work;
}

// One wave64 per 16-px tile, the tile's work key recorded by the tile
// (the batch's and the single-view default rasterizer).
template <bool COUNT>
__device__ inline void rasterize_tile(uint32_t tx, uint32_t ty, uint32_t vtile, uint32_t start, uint32_t end,
                                      const float4 *__restrict__ rec, const uint32_t *__restrict__ pvals,
                                      float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba,
                                      const RastParams &p, unsigned long long *__restrict__ d_iters,
                                      RastLds &L) {
  const uint32_t work = rasterize_tile16<COUNT, false>(tx, ty, vtile, start, end, rec, pvals, accum, out_rgba, p,
                                                       d_iters, L);
  if (p.cost_key && threadIdx.x == 0) p.cost_key[vtile] = work_key(work);
}

// One wave64 per 8-px tile (tx, ty): lane l owns pixel (l % 8, l / 8) of the
// tile; strip h = 0, 1 holds rows 4h .. 4h + 3 (lanes 32h .. 32h + 31).
// Same staging, tests and loop body as the 16-px tile, with closed flags
// and skips per 8-px tile. Returns the work measure.
template <bool MULTI>
__device__ inline uint32_t rasterize_tile8(uint32_t tx, uint32_t ty, uint32_t vtile, uint32_t start, uint32_t end,
                                           const float4 *__restrict__ rec, const uint32_t *__restrict__ pvals,
                                           float4 *__restrict__ accum, uint32_t *__restrict__ out_rgba,
                                           const RastParams &p, RastLds &L) {
  if (tile_skipped(p, vtile, start, end)) return 0;
  const int lane = threadIdx.x & 63;
  const uint32_t px = tx * 8u + (uint32_t)(lane & 7), py = ty * 8u + (uint32_t)(lane >> 3);
  const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
  const bool inb = px < p.width && py < p.height;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = inb ? 1.0f : 0.0f;
  if (inb && p.resume) {
    const float4 a = accum[(size_t)py * p.width + px];
    cr = a.x; cg = a.y; cb = a.z; T = a.w;
  }
  uint32_t work = 0;
  if (start < end) {
    const float tile_x0 = (float)(tx * 8u) + 0.5f, tile_x1 = tile_x0 + 7.0f;
    const uint32_t mine = (uint32_t)lane >> 5;  // this lane's strip
    for (uint32_t base = start; base < end; base += kStage) {
      const uint64_t live = __ballot(!(T < kMinT));
      if (live == 0) break;
      // lanes h * 32 + c stage splat c and test it against strip h
      const uint32_t slot = (uint32_t)lane & (kStage - 1u), half = (uint32_t)lane / kStage;
      const uint32_t i = base + slot;
      bool hit = false;
      if (i < end) {
        const uint32_t j = pvals[i];
        const float4 a = rec[3 * (size_t)j + 0];
        const float4 b = rec[3 * (size_t)j + 1];
        const float4 col = rec[3 * (size_t)j + 2];
        if (half == 0) {
          L.rec[3 * slot + 0] = a;
          L.rec[3 * slot + 1] = b;
          L.rec[3 * slot + 2] = col;
        }
        const float y0 = (float)(ty * 8u + 4u * half) + 0.5f;
        hit = strip_may_hit(strip_test(a.x, a.y, a.z, b.x, b.y, b.z, col.w, tile_x0, tile_x1), y0, y0 + 3.0f);
      }
      const uint64_t bits = __ballot(hit);
      const uint32_t s0 = (uint32_t)live ? (uint32_t)bits : 0u;
      const uint32_t s1 = (uint32_t)(live >> 32) ? (uint32_t)(bits >> 32) : 0u;
      work += 8u + (uint32_t)(__builtin_popcount(s0) + __builtin_popcount(s1));
      stage_sync<MULTI>();
      const uint32_t my_bits = mine ? s1 : s0;
      uint32_t todo = s0 | s1;
      while (todo) {
        const uint32_t c = (uint32_t)__builtin_ctz(todo);
        todo &= todo - 1u;
        if (!((my_bits >> c) & 1u)) continue;
        const float4 a = L.rec[3 * c + 0], b = L.rec[3 * c + 1];
        blend_row(col_terms(a, b, pxf), a, b, L.rec[3 * c + 2], pyf, cr, cg, cb, T);
      }
      stage_sync<MULTI>();
    }
  }
  bool closing = false;
  if (p.closed && !p.finalize) {
    const bool open = __ballot(!(T < kMinT)) != 0;
    if (lane == 0) p.closed[vtile] = open ? 0 : 1;
    closing = !open && p.final_on_close;
  }
  if (inb) {
    const size_t pix = (size_t)py * p.width + px;
    if (p.write_accum && !closing) accum[pix] = make_float4(cr, cg, cb, T);
    if (p.finalize || closing) out_rgba[pix] = pack_rgba(cr, cg, cb, T, p);
  }
  return work;
}

// Open-tile map for the next pass: per row, prefix counts of open tiles.
// One 256-thread block per row (cols <= 256).
__global__ __launch_bounds__(256) void k_open_prefix(const uint8_t *__restrict__ closed, uint32_t cols,
                                                     uint16_t *__restrict__ pre) {
  __shared__ uint32_t s_w[4];
  const uint32_t y = blockIdx.x, x = threadIdx.x;
  const uint32_t open = (x < cols && !closed[(size_t)y * cols + x]) ? 1u : 0u;
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

__global__ void k_fill_u8(uint8_t *__restrict__ p, size_t n, uint8_t v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace
}  // namespace dl
