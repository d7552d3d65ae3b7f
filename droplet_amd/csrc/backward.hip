// Backward of the forward splat renderer on gfx950 (no reference counterpart:
// droplet renders only; README.md:106). Differentiates the forward exactly
// as render.hip evaluates it: the rasterizer loop of rasterize.ts:62-79 and
// the EWA projection of project.ts:59-234, quirks included.
//
//   k_image_rgb          the differentiable image: clamp(C + T * bg, 0, 1)
//                        per channel (finalize.ts:31-35 before quantisation)
//   k_backward_raster    one wave per tile. Each pixel re-runs its list front
//                        to back, recomputing T instead of storing it, and
//                        takes dL/d(sx, sy, A, B, C, a, r, g, b) of every
//                        splat it composited. The wave sums over its 256
//                        pixels and adds once per (tile, splat).
//   k_backward_project   per splat: chain rule from those 9 values to the
//                        position, rotation, log-scale, opacity logit, DC
//                        and SH coefficients, recomputing the projection.
//
// Compositing: out_c = sum_i col_ic a_i T_i + T_end bg_c with T_{i+1} =
// T_i (1 - a_i). d out_c / d col_ic = a_i T_i and d out_c / d a_i =
// col_ic T_i - S_ic / (1 - a_i), where S_ic = out_c - sum_{j<=i} col_jc a_j
// T_j is what lies behind splat i (including the background).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "dl_common.h"
#include "dl_internal.h"
#include "render.h"
#include "render_device.h"

namespace dl {

namespace {

constexpr int kSplatGrads = 9;  // sx, sy, A, B, C, a, r, g, b

__global__ __launch_bounds__(256) void k_image_rgb(const float4 *__restrict__ accum, size_t npx, float bg_r,
                                                   float bg_g, float bg_b, float *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npx) return;
  const float4 a = accum[i];
  out[3 * i + 0] = wclamp(a.x + a.w * bg_r, 0.0f, 1.0f);
  out[3 * i + 1] = wclamp(a.y + a.w * bg_g, 0.0f, 1.0f);
  out[3 * i + 2] = wclamp(a.z + a.w * bg_b, 0.0f, 1.0f);
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ inline float wave_sum_f(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(64) void k_backward_raster(const float4 *__restrict__ rec,
                                                        const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ pvals,
                                                        const float4 *__restrict__ accum,
                                                        const float *__restrict__ grad_rgb, RastParams p,
                                                        float *__restrict__ gsplat) {
  __shared__ float4 s_rec[64 * 3];
  __shared__ uint32_t s_id[64];
  __shared__ uint32_t s_mask[64];
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % p.tiles_x, ty = tile / p.tiles_x;
  const int lane = threadIdx.x;
  const uint32_t px = tx * kTile + (lane & 15);
  const uint32_t py0 = ty * kTile + (lane >> 4);
  const float pxf = (float)px + 0.5f;
  const float bg[3] = {p.bg_r, p.bg_g, p.bg_b};
  float T[4], pyf[4], G[4][3], S[4][3];
  bool live[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t y = py0 + 4 * k;
    const bool inb = px < p.width && y < p.height;
    pyf[k] = (float)y + 0.5f;
    T[k] = 1.0f;
    live[k] = inb;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      G[k][c] = 0.0f;
      S[k][c] = 0.0f;
    }
    if (inb) {
      const size_t pix = (size_t)y * p.width + px;
      const float4 a = accum[pix];
      const float fin[3] = {a.x, a.y, a.z};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        // clamp(x, 0, 1) passes the gradient only strictly inside (0, 1)
        const float pre = fin[c] + a.w * bg[c];
        G[k][c] = (pre > 0.0f && pre < 1.0f) ? grad_rgb[3 * pix + c] : 0.0f;
        S[k][c] = pre;  // what lies behind the next splat: starts as the whole output
      }
    }
  }
  const uint32_t start = off[tile], end = off[tile + 1];
  const float tile_x0 = (float)(tx * kTile) + 0.5f, tile_x1 = tile_x0 + 15.0f;
  for (uint32_t base = start; base < end; base += 64) {
    uint32_t strips_live = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) strips_live |= (__ballot(live[k]) != 0 ? 1u : 0u) << k;
    if (strips_live == 0) break;
    const uint32_t i = base + lane;
    if (i < end) {
      const uint32_t j = pvals[i];
      const float4 a = rec[3 * (size_t)j + 0];
      const float4 b = rec[3 * (size_t)j + 1];
      const float4 col = rec[3 * (size_t)j + 2];
      s_rec[3 * lane + 0] = a;
      s_rec[3 * lane + 1] = b;
      s_rec[3 * lane + 2] = col;
      s_id[lane] = j;
      uint32_t m = 0;
      const StripTest st = strip_test(a.x, a.y, a.z, b.x, b.y, b.z, col.w, tile_x0, tile_x1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float y0 = (float)(ty * kTile + 4 * k) + 0.5f;
        if (strip_may_hit(st, y0, y0 + 3.0f)) m |= 1u << k;
      }
      s_mask[lane] = m;
    }
    __syncthreads();
    const uint32_t cnt = min(64u, end - base);
    for (uint32_t c = 0; c < cnt; c++) {
      const uint32_t m = __builtin_amdgcn_readfirstlane(s_mask[c]) & strips_live;
      if (m == 0) continue;
      const float4 a = s_rec[3 * c + 0];
      const float4 b = s_rec[3 * c + 1];
      const float4 col = s_rec[3 * c + 2];
      float g[kSplatGrads];
#pragma unroll
      for (int q = 0; q < kSplatGrads; q++) g[q] = 0.0f;
      bool any = false;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!((m >> k) & 1u)) continue;
        if (!live[k]) continue;
        const float dx = pxf - a.x;
        const float dy = pyf[k] - a.y;
        const float r = a.z;
        if (r <= 0.0f || __builtin_fabsf(dx) > r || __builtin_fabsf(dy) > r) continue;
        const float power = gauss_power(b.x, b.y, b.z, dx, dy);
        if (power > 0.0f || power < col.w) continue;
        const float e = dl_expf_nonpos(power);
        const float raw = b.w * wmax(0.0f, e - kGaussFloor);
        const float alpha = wmin(kOpacityCap, raw);
        if (alpha < kMinAlpha) continue;
        any = true;
        const float w = T[k] * alpha;
        const float cc[3] = {col.x, col.y, col.z};
        float dalpha = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          S[k][q] = S[k][q] - cc[q] * w;  // now: everything behind this splat
          g[6 + q] += G[k][q] * w;
          dalpha += G[k][q] * (cc[q] * T[k] - S[k][q] / (1.0f - alpha));
        }
        if (raw < kOpacityCap) {
          // alpha = a * (e - floor), e = exp(power)
          g[5] += dalpha * (e - kGaussFloor);
          const float dpow = dalpha * b.w * e;
          g[0] += dpow * (b.x * dx + b.y * dy);
          g[1] += dpow * (b.y * dx + b.z * dy);
          g[2] += dpow * (-0.5f * dx * dx);
          g[3] += dpow * (-dx * dy);
          g[4] += dpow * (-0.5f * dy * dy);
        }
        T[k] = T[k] * (1.0f - alpha);
        live[k] = !(T[k] < kMinT);
      }
      if (__ballot(any) == 0) continue;
#pragma unroll
      for (int q = 0; q < kSplatGrads; q++) g[q] = wave_sum_f(g[q]);
      if (lane == 0) {
        float *dst = gsplat + (size_t)kSplatGrads * s_id[c];
#pragma unroll
        for (int q = 0; q < kSplatGrads; q++)
          if (g[q] != 0.0f) atomicAdd(dst + q, g[q]);
      }
    }
    __syncthreads();
  }
}

// d basis / d (x, y, z) of the SH terms of bands 1..3 in eval_sh's order
// (coefficient k of a channel), accumulated as gd += w_k * grad b_k.
template <int BANDS, class R>
__device__ inline void sh_dir_grad(const R *w, R x, R y, R z, R gd[3]) {
  const R SH1 = (R)(float)0.4886025119029199;
  gd[1] += w[0] * (-SH1);
  gd[2] += w[1] * SH1;
  gd[0] += w[2] * (-SH1);
  if constexpr (BANDS >= 2) {
    const R K0 = (R)(float)1.0925484305920792, K2 = (R)(float)0.31539156525252005, K4 = (R)(float)0.5462742152960396;
    gd[0] += w[3] * (K0 * y);
    gd[1] += w[3] * (K0 * x);
    gd[1] += w[4] * (-K0 * z);
    gd[2] += w[4] * (-K0 * y);
    gd[0] += w[5] * (-2.0f * K2 * x);
    gd[1] += w[5] * (-2.0f * K2 * y);
    gd[2] += w[5] * (4.0f * K2 * z);
    gd[0] += w[6] * (-K0 * z);
    gd[2] += w[6] * (-K0 * x);
    gd[0] += w[7] * (2.0f * K4 * x);
    gd[1] += w[7] * (-2.0f * K4 * y);
    if constexpr (BANDS >= 3) {
      const R Q0 = (R)(float)-0.5900435899266435, Q1 = (R)(float)2.890611442640554;
      const R Q2 = (R)(float)-0.4570457994644658, Q3 = (R)(float)0.3731763325901154, Q5 = (R)(float)1.445305721320277;
      const R xx = x * x, yy = y * y, zz = z * z;
      gd[0] += w[8] * (Q0 * 6.0f * x * y);
      gd[1] += w[8] * (Q0 * (3.0f * xx - 3.0f * yy));
      gd[0] += w[9] * (Q1 * y * z);
      gd[1] += w[9] * (Q1 * x * z);
      gd[2] += w[9] * (Q1 * x * y);
      gd[0] += w[10] * (Q2 * y * (-2.0f * x));
      gd[1] += w[10] * (Q2 * (4.0f * zz - xx - 3.0f * yy));
      gd[2] += w[10] * (Q2 * y * 8.0f * z);
      gd[0] += w[11] * (Q3 * z * (-6.0f * x));
      gd[1] += w[11] * (Q3 * z * (-6.0f * y));
      gd[2] += w[11] * (Q3 * (6.0f * zz - 3.0f * xx - 3.0f * yy));
      gd[0] += w[12] * (Q2 * (4.0f * zz - 3.0f * xx - yy));
      gd[1] += w[12] * (Q2 * x * (-2.0f * y));
      gd[2] += w[12] * (Q2 * x * 8.0f * z);
      gd[0] += w[13] * (Q5 * z * 2.0f * x);
      gd[1] += w[13] * (-Q5 * z * 2.0f * y);
      gd[2] += w[13] * (Q5 * (xx - yy));
      gd[0] += w[14] * (Q0 * (3.0f * xx - 3.0f * yy));
      gd[1] += w[14] * (Q0 * x * (-6.0f * y));
    }
  }
}

// SH basis values in eval_sh's coefficient order
template <int BANDS, class R>
__device__ inline void sh_basis(R x, R y, R z, R *bv) {
  const R SH1 = (R)(float)0.4886025119029199;
  bv[0] = (-SH1) * y;
  bv[1] = SH1 * z;
  bv[2] = (-SH1) * x;
  if constexpr (BANDS >= 2) {
    const R xx = x * x, yy = y * y, zz = z * z;
    bv[3] = (R)(float)1.0925484305920792 * (x * y);
    bv[4] = (R)(float)-1.0925484305920792 * (y * z);
    bv[5] = (R)(float)0.31539156525252005 * ((2 * zz - xx) - yy);
    bv[6] = (R)(float)-1.0925484305920792 * (x * z);
    bv[7] = (R)(float)0.5462742152960396 * (xx - yy);
    if constexpr (BANDS >= 3) {
      bv[8] = ((R)(float)-0.5900435899266435 * y) * (3 * xx - yy);
      bv[9] = ((R)(float)2.890611442640554 * (x * y)) * z;
      bv[10] = ((R)(float)-0.4570457994644658 * y) * ((4 * zz - xx) - yy);
      bv[11] = ((R)(float)0.3731763325901154 * z) * ((2 * zz - 3 * xx) - 3 * yy);
      bv[12] = ((R)(float)-0.4570457994644658 * x) * ((4 * zz - xx) - yy);
      bv[13] = ((R)(float)1.445305721320277 * z) * (xx - yy);
      bv[14] = ((R)(float)-0.5900435899266435 * x) * (xx - 3 * yy);
    }
  }
}

// Gradient of o = quat_rot(q, v) (project.ts: o = v + w t + qv x t, t = 2 qv
// x v, for any q, normalised or not) with respect to q, given go = dL/do.
template <class R>
__device__ inline void quat_rot_grad(const R q[4], const R v[3], const R go[3], R gq[4]) {
  const R qv[3] = {q[1], q[2], q[3]};
  const R cx[3] = {qv[1] * v[2] - qv[2] * v[1], qv[2] * v[0] - qv[0] * v[2], qv[0] * v[1] - qv[1] * v[0]};
  // o = v + 2 w (qv x v) + 2 qv (qv . v) - 2 v |qv|^2
  gq[0] += 2 * (go[0] * cx[0] + go[1] * cx[1] + go[2] * cx[2]);
  const R vxg[3] = {v[1] * go[2] - v[2] * go[1], v[2] * go[0] - v[0] * go[2], v[0] * go[1] - v[1] * go[0]};
  const R qdv = qv[0] * v[0] + qv[1] * v[1] + qv[2] * v[2];
  const R gdq = go[0] * qv[0] + go[1] * qv[1] + go[2] * qv[2];
  const R gdv = go[0] * v[0] + go[1] * v[1] + go[2] * v[2];
#pragma unroll
  for (int k = 0; k < 3; k++) gq[1 + k] += 2 * q[0] * vxg[k] + 2 * (go[k] * qdv + v[k] * gdq) - 4 * gdv * qv[k];
}

// quat_rot (project.ts:50-53) for any real type
template <class R>
__device__ inline void quat_rot_r(const R q[4], const R v[3], R o[3]) {
  const R c0 = q[2] * v[2] - q[3] * v[1], c1 = q[3] * v[0] - q[1] * v[2], c2 = q[1] * v[1] - q[2] * v[0];
  const R t0 = 2 * c0, t1 = 2 * c1, t2 = 2 * c2;
  o[0] = (v[0] + q[0] * t0) + (q[2] * t2 - q[3] * t1);
  o[1] = (v[1] + q[0] * t1) + (q[3] * t0 - q[1] * t2);
  o[2] = (v[2] + q[0] * t2) + (q[1] * t1 - q[2] * t0);
}

// The chain rule of project.ts:59-234 per splat. The forward is f32, but
// near-singular projected covariances make f32 cancellation (det_orig =
// cov00 cov11 - cov01^2, the antialiasing factor) matter at the 1e-3 level.
// So the derivative of the projection is taken in f64, from the same f32
// inputs.
template <int BANDS>
__global__ __launch_bounds__(256) void k_backward_project(SceneView sc, ProjUniforms u,
                                                          const float4 *__restrict__ rec,
                                                          const float *__restrict__ gsplat,
                                                          float *__restrict__ gpos, float *__restrict__ grot,
                                                          float *__restrict__ gls, float *__restrict__ glogit,
                                                          float *__restrict__ gdc, float *__restrict__ gsh) {
  using R = double;
  constexpr int M = BANDS == 0 ? 0 : BANDS == 1 ? 3 : BANDS == 2 ? 8 : 15;
  const uint32_t id = blockIdx.x * 256 + threadIdx.x;
  if (id >= sc.count) return;
  R gp[3] = {0, 0, 0}, gq[4] = {0, 0, 0, 0}, gl[3] = {0, 0, 0}, go = 0, gd[3] = {0, 0, 0};
  R gshv[3 * (M > 0 ? M : 1)];
#pragma unroll
  for (int k = 0; k < 3 * (M > 0 ? M : 1); k++) gshv[k] = 0;
  const float *g9 = gsplat + (size_t)kSplatGrads * id;
  R g[kSplatGrads];
  bool any = false;
#pragma unroll
  for (int q = 0; q < kSplatGrads; q++) {
    g[q] = g9[q];
    any = any || g9[q] != 0.0f;
  }
  const float4 r0 = rec[3 * (size_t)id];
  if (any && r0.z > 0.0f) {
    // ---- recompute the projection ----
    const float *m = u.view;
    const R p3[3] = {sc.pos[3 * (size_t)id], sc.pos[3 * (size_t)id + 1], sc.pos[3 * (size_t)id + 2]};
    const float4 qv4 = reinterpret_cast<const float4 *>(sc.rot)[id];
    const R q[4] = {qv4.x, qv4.y, qv4.z, qv4.w};
    const R s[3] = {exp((R)sc.ls[3 * (size_t)id]), exp((R)sc.ls[3 * (size_t)id + 1]),
                    exp((R)sc.ls[3 * (size_t)id + 2])};
    const R logit = sc.op[id];
    const R W[3][3] = {{m[0], m[4], m[8]}, {m[1], m[5], m[9]}, {m[2], m[6], m[10]}};
    const R pcx = W[0][0] * p3[0] + W[0][1] * p3[1] + W[0][2] * p3[2] + (R)m[12];
    const R pcy = W[1][0] * p3[0] + W[1][1] * p3[1] + W[1][2] * p3[2] + (R)m[13];
    const R pcz = W[2][0] * p3[0] + W[2][1] * p3[1] + W[2][2] * p3[2] + (R)m[14];
    const R f = u.focal;
    R Rm[3][3];  // columns: quat_rot(q, e_j)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      R e[3] = {0, 0, 0};
      e[j] = 1;
      R o[3];
      quat_rot_r(q, e, o);
      Rm[0][j] = o[0];
      Rm[1][j] = o[1];
      Rm[2][j] = o[2];
    }
    R Mx[3][3];  // M = R diag(s)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int j = 0; j < 3; j++) Mx[a][j] = Rm[a][j] * s[j];
    R C3[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) C3[a][b] = Mx[a][0] * Mx[b][0] + Mx[a][1] * Mx[b][1] + Mx[a][2] * Mx[b][2];
    R V[3][3];  // W C3 W^T
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        R acc = 0;
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
          for (int l = 0; l < 3; l++) acc += W[a][k] * C3[k][l] * W[b][l];
        V[a][b] = acc;
      }
    const R iz = 1 / pcz;
    const R j00 = f * iz, j02 = -f * pcx * iz * iz, j11 = f * iz, j12 = -f * pcy * iz * iz;
    const R cov00 = j00 * j00 * V[0][0] + 2 * j00 * j02 * V[0][2] + j02 * j02 * V[2][2];
    const R cov01 = j00 * j11 * V[0][1] + j00 * j12 * V[0][2] + j02 * j11 * V[1][2] + j02 * j12 * V[2][2];
    const R cov11 = j11 * j11 * V[1][1] + 2 * j11 * j12 * V[1][2] + j12 * j12 * V[2][2];
    const R det_orig = cov00 * cov11 - cov01 * cov01;
    const R kaa = (R)kAA;
    const R d00 = cov00 + kaa, d11 = cov11 + kaa;
    const R det = d00 * d11 - cov01 * cov01;
    const R ratio = det_orig / det;
    const R aa = ratio > 0 ? sqrt(ratio) : 0;
    const R mid = 0.5 * (d00 + d11);
    const R disc = mid * mid - det;
    const R r_eig = sqrt(disc > (R)kEigMin ? disc : (R)kEigMin);
    const R lambda1 = mid + r_eig;
    const R radius_raw = (R)kSigma * sqrt(lambda1);
    const R fade_s = (R)kFadeStart * (R)u.height, fade_e = (R)kFadeEnd * (R)u.height;
    const R fade_arg = (fade_e - radius_raw) / (fade_e - fade_s);
    const R fade = fade_arg < 0 ? 0 : fade_arg > 1 ? 1 : fade_arg;
    const R sig = 1 / (1 + exp(-logit));
    // ---- colour ----
    R len = 1, dir[3] = {0, 0, 0};
    R bv[M > 0 ? M : 1];
    R colpre[3] = {sc.dc[3 * (size_t)id], sc.dc[3 * (size_t)id + 1], sc.dc[3 * (size_t)id + 2]};
    if constexpr (BANDS > 0) {
      len = sqrt(pcx * pcx + pcy * pcy + pcz * pcz);
      dir[0] = pcx / len;
      dir[1] = pcy / len;
      dir[2] = pcz / len;
      sh_basis<BANDS, R>(dir[0], dir[1], dir[2], bv);
      const float *shc = sc.sh + (size_t)id * 3 * M;
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < M; k++) colpre[c] += bv[k] * (R)shc[c * M + k];
    }
    // ---- backward ----
    R gpc[3] = {0, 0, 0};
    R gdir[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const R gc = (colpre[c] + 0.5 > 0) ? g[6 + c] : 0;  // max(col + 0.5, 0)
      gd[c] = gc;
      if constexpr (BANDS > 0) {
        const float *shc = sc.sh + (size_t)id * 3 * M;
        R w[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
          gshv[c * M + k] = gc * bv[k];
          w[k] = gc * (R)shc[c * M + k];
        }
        sh_dir_grad<BANDS, R>(w, dir[0], dir[1], dir[2], gdir);
      }
    }
    if constexpr (BANDS > 0) {
      // dir = pc / |pc|
      const R dd = gdir[0] * dir[0] + gdir[1] * dir[1] + gdir[2] * dir[2];
#pragma unroll
      for (int k = 0; k < 3; k++) gpc[k] += (gdir[k] - dd * dir[k]) / len;
    }
    // opacity a = (sigmoid(logit) * fade) * aa
    const R ga = g[5];
    go = ga * fade * aa * sig * (1 - sig);
    const R gaa = ga * sig * fade;
    const R gfade = ga * sig * aa;
    R glambda = 0;
    if (fade_arg > 0 && fade_arg < 1) {
      const R grr = gfade * (-1 / (fade_e - fade_s));
      glambda += grr * (R)kSigma * 0.5 / sqrt(lambda1);
    }
    // conic A = d11 / det, B = -cov01 / det, C = d00 / det
    R gd00 = g[4] / det, gd11 = g[2] / det, gc01 = -g[3] / det;
    R gdet = -(g[2] * d11 - g[3] * cov01 + g[4] * d00) / (det * det);
    R gdo = 0;
    if (ratio > 0) {
      const R gr = gaa * 0.5 / aa;  // d sqrt(ratio)
      gdo += gr / det;
      gdet += -gr * det_orig / (det * det);
    }
    // lambda1 = mid + sqrt(max(kEigMin, mid^2 - det))
    R gmid = glambda;
    if (disc > (R)kEigMin) {
      gmid += glambda * mid / r_eig;
      gdet += -glambda * 0.5 / r_eig;
    }
    gd00 += 0.5 * gmid;
    gd11 += 0.5 * gmid;
    // det = d00 d11 - cov01^2 ; det_orig = cov00 cov11 - cov01^2
    gd00 += gdet * d11;
    gd11 += gdet * d00;
    gc01 += -2 * gdet * cov01;
    const R gc00 = gd00 + gdo * cov11;
    const R gc11 = gd11 + gdo * cov00;
    gc01 += -2 * gdo * cov01;
    // cov2 = J V J^T over the unique entries cov00, cov01, cov11
    R gV[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    gV[0][0] = gc00 * j00 * j00;
    gV[1][1] = gc11 * j11 * j11;
    gV[2][2] = gc00 * j02 * j02 + gc01 * j02 * j12 + gc11 * j12 * j12;
    gV[0][1] = gc01 * j00 * j11;                            // unique-entry gradients
    gV[0][2] = 2 * gc00 * j00 * j02 + gc01 * j00 * j12;     // (each off-diagonal once)
    gV[1][2] = gc01 * j02 * j11 + 2 * gc11 * j11 * j12;
    const R gj00 = gc00 * (2 * j00 * V[0][0] + 2 * j02 * V[0][2]) + gc01 * (j11 * V[0][1] + j12 * V[0][2]);
    const R gj02 = gc00 * (2 * j00 * V[0][2] + 2 * j02 * V[2][2]) + gc01 * (j11 * V[1][2] + j12 * V[2][2]);
    const R gj11 = gc01 * (j00 * V[0][1] + j02 * V[1][2]) + gc11 * (2 * j11 * V[1][1] + 2 * j12 * V[1][2]);
    const R gj12 = gc01 * (j00 * V[0][2] + j02 * V[2][2]) + gc11 * (2 * j11 * V[1][2] + 2 * j12 * V[2][2]);
    // J(pc): j00 = j11 = f/z, j02 = -f x / z^2, j12 = -f y / z^2
    gpc[0] += gj02 * (-f * iz * iz);
    gpc[1] += gj12 * (-f * iz * iz);
    gpc[2] += (gj00 + gj11) * (-f * iz * iz) + gj02 * (2 * f * pcx * iz * iz * iz) +
              gj12 * (2 * f * pcy * iz * iz * iz);
    // screen position: sx = -f x / z + cx, sy = f y / z + cy
    gpc[0] += g[0] * (-f * iz);
    gpc[2] += g[0] * (f * pcx * iz * iz);
    gpc[1] += g[1] * (f * iz);
    gpc[2] += g[1] * (-f * pcy * iz * iz);
    // V = W C3 W^T: symmetric gradient Gh (off-diagonal halves), dC3 = W^T Gh W
    R Gh[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        Gh[a][b] = (a == b) ? gV[a][a] : 0.5 * gV[lo][hi];
      }
    R GC[3][3];  // dL/dC3 as a full symmetric matrix
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int l = 0; l < 3; l++) {
        R acc = 0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int b = 0; b < 3; b++) acc += W[a][k] * Gh[a][b] * W[b][l];
        GC[k][l] = acc;
      }
    // C3 = M M^T: dL/dM = 2 GC M ; M[:, j] = R e_j * s_j
    R gs[3] = {0, 0, 0};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      R gm[3];
#pragma unroll
      for (int a = 0; a < 3; a++) gm[a] = 2 * (GC[a][0] * Mx[0][j] + GC[a][1] * Mx[1][j] + GC[a][2] * Mx[2][j]);
      gs[j] = gm[0] * Rm[0][j] + gm[1] * Rm[1][j] + gm[2] * Rm[2][j];
      const R gr[3] = {gm[0] * s[j], gm[1] * s[j], gm[2] * s[j]};
      R e[3] = {0, 0, 0};
      e[j] = 1;
      quat_rot_grad<R>(q, e, gr, gq);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) gl[j] = gs[j] * s[j];
    // pc = W p + t
#pragma unroll
    for (int k = 0; k < 3; k++) gp[k] = W[0][k] * gpc[0] + W[1][k] * gpc[1] + W[2][k] * gpc[2];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    gpos[3 * (size_t)id + k] = (float)gp[k];
    gls[3 * (size_t)id + k] = (float)gl[k];
    gdc[3 * (size_t)id + k] = (float)gd[k];
  }
  reinterpret_cast<float4 *>(grot)[id] = make_float4((float)gq[0], (float)gq[1], (float)gq[2], (float)gq[3]);
  glogit[id] = (float)go;
  if constexpr (M > 0) {
#pragma unroll
    for (int k = 0; k < 3 * M; k++) gsh[(size_t)id * 3 * M + k] = (float)gshv[k];
  }
}

}  // namespace

void Renderer::forward_train(const SceneView &sc, uint64_t scene_serial, const dl_camera &cam,
                             const dl_background &bg, float *out_rgb) {
  if (cam.projection != DL_PINHOLE)
    fail(DL_ERR_INVALID, "the differentiable render supports pinhole cameras (equirect has no backward)");
  train_mode = true;
  try {
    render(sc, cam, bg);
  } catch (...) {
    train_mode = false;
    train_valid = false;
    throw;
  }
  train_mode = false;
  train_valid = true;
  train_scene = scene_serial;
  train_sc = sc;
  train_u = make_uniforms(cam, max_cov);
  train_bg = bg;
  hipStream_t s = ctx->stream;
  const size_t npx = (size_t)width * height;
  img_rgb.reserve(npx * 12);
  k_image_rgb<<<div_up(npx, 256), 256, 0, s>>>(accum.as<float4>(), npx, bg.r, bg.g, bg.b, img_rgb.as<float>());
  DL_HIP(hipGetLastError());
  if (out_rgb) ctx->staging.to_host(out_rgb, img_rgb.get(), npx * 12, s);
  DL_HIP(hipStreamSynchronize(s));
}

void Renderer::backward(const float *grad_rgb, const dl_grads &out) {
  if (!train_valid)
    fail(DL_ERR_INVALID, "dl_render_backward needs a dl_render_forward with no other render of this renderer since");
  if (!scene_alive(train_scene)) fail(DL_ERR_INVALID, "the scene of the last dl_render_forward was destroyed");
  const SceneView &sc = train_sc;
  const uint32_t n = sc.count;
  const size_t npx = (size_t)width * height;
  const uint32_t M = sc.bands == 0 ? 0 : sc.bands == 1 ? 3 : sc.bands == 2 ? 8 : 15;
  hipStream_t s = ctx->stream;
  ctx->make_current();
  grad_img.reserve(std::max<size_t>(npx, 1) * 12);
  ctx->staging.to_device(grad_img.get(), grad_rgb, npx * 12, s);
  const size_t nn = std::max<uint32_t>(n, 1);
  gsplat.reserve(nn * kSplatGrads * 4);
  gparams.reserve(nn * (3 + 4 + 3 + 1 + 3 + 3 * (size_t)M) * 4 + 1024);
  float *gpos = gparams.as<float>();
  float *grot = gpos + 3 * nn;
  float *gls = grot + 4 * nn;
  float *glogit = gls + 3 * nn;
  float *gdc = glogit + nn;
  float *gsh = gdc + 3 * nn;
  if (n) {
    k_zero_u32<<<div_up((size_t)n * kSplatGrads, 256), 256, 0, s>>>(gsplat.as<uint32_t>(), (size_t)n * kSplatGrads);
    DL_HIP(hipGetLastError());
    if (stats.num_visible > 0 && stats.num_pairs > 0) {
      RastParams p{};
      p.tiles_x = stats.tiles_x;
      p.width = width;
      p.height = height;
      p.bg_r = train_bg.r;
      p.bg_g = train_bg.g;
      p.bg_b = train_bg.b;
      p.bg_a = train_bg.a;
      k_backward_raster<<<stats.num_tiles, 64, 0, s>>>(rec.as<float4>(), ranges.as<uint32_t>(), train_vals, accum.as<float4>(),
                                                      grad_img.as<float>(), p, gsplat.as<float>());
      DL_HIP(hipGetLastError());
    }
    const uint32_t blocks = div_up(n, 256);
    switch (sc.bands) {
      case 0: k_backward_project<0><<<blocks, 256, 0, s>>>(sc, train_u, rec.as<float4>(), gsplat.as<float>(), gpos, grot, gls, glogit, gdc, gsh); break;
      case 1: k_backward_project<1><<<blocks, 256, 0, s>>>(sc, train_u, rec.as<float4>(), gsplat.as<float>(), gpos, grot, gls, glogit, gdc, gsh); break;
      case 2: k_backward_project<2><<<blocks, 256, 0, s>>>(sc, train_u, rec.as<float4>(), gsplat.as<float>(), gpos, grot, gls, glogit, gdc, gsh); break;
      default: k_backward_project<3><<<blocks, 256, 0, s>>>(sc, train_u, rec.as<float4>(), gsplat.as<float>(), gpos, grot, gls, glogit, gdc, gsh); break;
    }
    DL_HIP(hipGetLastError());
  }
  auto down = [&](float *dst, const float *src, size_t count) {
    if (dst && count) ctx->staging.to_host(dst, src, count * 4, s);
  };
  down(out.positions, gpos, 3 * (size_t)n);
  down(out.rotations, grot, 4 * (size_t)n);
  down(out.log_scales, gls, 3 * (size_t)n);
  down(out.opacity_logits, glogit, n);
  down(out.colors_dc, gdc, 3 * (size_t)n);
  down(out.colors_sh, gsh, 3 * (size_t)M * n);
  DL_HIP(hipStreamSynchronize(s));
}

}  // namespace dl
