// Shared device helpers for the gfx950 kernels of libdroplet_hip.so.
//
// Numerics contract. Every kernel is compiled with -ffp-contract=off: each
// f32/f64 operation rounds on its own, in the order the reference's WGSL/JS
// writes it (left to right). IEEE +,-,*,/,sqrt are correctly rounded on both
// gfx950 and the host, so those results are bit-identical to any other IEEE
// evaluation of the same expression. The only transcendental on the hot path,
// e^x, is fixed by one specification (dl_expf below), so keys, ranges, masks
// and images do not depend on an implementation-defined exp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dl {

// ---------------------------------------------------------------------------
// WGSL prelude constants (src/render/shaders/constants.ts:19-51). Each is the
// JS number printed into the shader text, parsed as an abstract float and
// converted to f32: (float)<double literal>.
// ---------------------------------------------------------------------------
constexpr uint32_t kTile = 16;
constexpr float kSigma = (float)3.0;
constexpr float kAA = (float)0.3;
constexpr float kOpacityCap = (float)0.999;
constexpr float kMinT = (float)0.0001;
constexpr float kMinAlpha = (float)0.00392156862745098;
constexpr float kGaussFloor = (float)1.1108996538e-2;  // Math.exp(-4.5).toExponential(10)
constexpr float kFadeStart = (float)0.9481481481481482;
constexpr float kFadeEnd = (float)1.8962962962962964;
constexpr float kDetMin = (float)1.0e-6;               // project.ts:165
constexpr float kEigMin = (float)0.1;                  // project.ts:181
constexpr float kTwoPi = (float)(3.14159265 * 2.0);    // project.ts:260 (abstract const-expr)
constexpr float kPi = (float)3.14159265;               // project.ts:261
constexpr uint32_t kMaxCoverageDefault = 16384;        // pipeline.ts:41
constexpr uint32_t kInvisibleKey = 0xFFFFFFFFu;        // sorts after every finite depth

__host__ __device__ inline float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// e^x: Cody-Waite reduction x = k*ln2 + r (ln2 = f32 head + f32 tail), the
// degree-7 Taylor polynomial of e^r on |r| <= ln2/2 (coefficients 1/n! in f32)
// by Horner with fused multiply-add, then 2^k applied as two exact powers of
// two so that a subnormal result rounds once. Faithful (< 1 ulp).
__host__ __device__ inline float dl_expf(float x) {
  if (x != x) return x;
  x = x < -104.0f ? -104.0f : x;
  x = x > 89.0f ? 89.0f : x;
  float kf = __builtin_rintf(x * 0x1.715476p+0f);
  float r = __builtin_fmaf(kf, -0x1.62e430p-1f, x);
  r = __builtin_fmaf(kf, 0x1.05c610p-29f, r);
  float p = 0x1.a01a02p-13f;
  p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
  p = __builtin_fmaf(p, r, 0x1.111112p-7f);
  p = __builtin_fmaf(p, r, 0x1.555556p-5f);
  p = __builtin_fmaf(p, r, 0x1.555556p-3f);
  p = __builtin_fmaf(p, r, 0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  int k = (int)kf;
  int k1 = k / 2;
  int k2 = k - k1;
  float s1 = bits_f32((uint32_t)(k1 + 127) << 23);
  float s2 = bits_f32((uint32_t)(k2 + 127) << 23);
  return (p * s1) * s2;
}

// dl_expf for x in [-87, 0] or NaN, bit-identical there. The clamps are no-ops
// on that range. 2^k (-126 <= k <= 0) is a normal power of two, so one exact
// scaling by ldexp equals the two exact steps. NaN propagates through p.
// The rasterizer's exponent always lies in [tau, 0], and tau > -4.6.
__host__ __device__ inline float dl_expf_nonpos(float x) {
  float kf = __builtin_rintf(x * 0x1.715476p+0f);
  float r = __builtin_fmaf(kf, -0x1.62e430p-1f, x);
  r = __builtin_fmaf(kf, 0x1.05c610p-29f, r);
  float p = 0x1.a01a02p-13f;
  p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
  p = __builtin_fmaf(p, r, 0x1.111112p-7f);
  p = __builtin_fmaf(p, r, 0x1.555556p-5f);
  p = __builtin_fmaf(p, r, 0x1.555556p-3f);
  p = __builtin_fmaf(p, r, 0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  // k from the bits of kf + 1.5 * 2^23 (exact for integral |kf| < 2^22);
  // defined for NaN too, where the result is NaN whatever k is
  const int k = (int)(f32_bits(kf + 0x1.8p23f) - 0x4B400000u);
  return __builtin_ldexpf(p, k);
}

// sigmoid (project.ts:55, voxel/shaders/voxelize.ts:82)
__host__ __device__ inline float dl_sigmoidf(float x) { return 1.0f / (1.0f + dl_expf(-x)); }

// WGSL u32(f32): round toward zero, clamp to [0, 2^32-1]; NaN -> 0.
__host__ __device__ inline uint32_t u32sat(float f) {
  if (!(f == f)) return 0u;
  if (f <= 0.0f) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}

// WGSL min/max/clamp on f32 (no NaN inputs reach them on the render path)
__host__ __device__ inline float wmax(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline float wmin(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }

// floatToSortable (src/spatial/radix-sort.ts:20-24)
__host__ __device__ inline uint32_t sortable_key(float f) {
  uint32_t u = f32_bits(f);
  return (u & 0x80000000u) ? ~u : (u ^ 0x80000000u);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

}  // namespace dl
