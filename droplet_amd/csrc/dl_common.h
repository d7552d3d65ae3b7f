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

// The sigma at which a voxel's Gaussian sum may stop early. The reference
// sums while sigma < 7 and calls a voxel solid when 1 - exp(-sigma) >=
// cutoff (voxel/shaders/voxelize.ts:111-134). Every term is >= 0, so sigma
// never decreases, and it never reaches 8 (the sum stops at 7, a term is at
// most 1); once it reaches the value returned here, 1 - dl_expf(-sigma) >=
// cutoff holds for every sigma up to 8, so the remaining terms cannot change
// the mask bit. The value is -ln((1 - c (1 + 2^-22)) / (1 + 2^-20)) rounded
// up to f32 and raised by two ulps: the margins exceed dl_expf's relative
// error and the rounding of 1 - e. It is 7 (the reference's own stop) for a
// cutoff outside (0, 1) or when the bound is not below 7.
inline float voxel_solid_stop(float cutoff) {
  const double c = cutoff;
  if (!(c > 0.0 && c < 1.0)) return 7.0f;
  const double e = (1.0 - c * (1.0 + 0x1p-22)) / (1.0 + 0x1p-20);
  if (!(e > 0.0)) return 7.0f;
  const double sd = -__builtin_log(e);
  float f = (float)sd;
  if ((double)f < sd) f = __builtin_nextafterf(f, 8.0f);
  f = __builtin_nextafterf(__builtin_nextafterf(f, 8.0f), 8.0f);
  return f < 7.0f ? f : 7.0f;
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

// Tile grid of a frame: the tile side in pixels (16 as in the reference,
// constants.ts:19, or the autotuner's 8 or 32), the grid at that size and
// the 16-px grid of the same image.
struct TileGrid {
  uint32_t tile, tiles_x, tiles_y, t16x, t16y;
};

__host__ __device__ inline TileGrid tile_grid(uint32_t width, uint32_t height, uint32_t tile) {
  return TileGrid{tile, (width + tile - 1) / tile, (height + tile - 1) / tile, (width + kTile - 1) / kTile,
                  (height + kTile - 1) / kTile};
}

// The reference's 16-px tile range of a splat (centre sx, sy, radius r),
// project.ts:223-233; empty when x1 < x0 or y1 < y0.
__host__ __device__ inline void tile_range16(float sx, float sy, float r, uint32_t t16x, uint32_t t16y,
                                             uint32_t &x0, uint32_t &x1, uint32_t &y0, uint32_t &y1) {
  x0 = u32sat(wmax(0.0f, __builtin_floorf((sx - r) / 16.0f)));
  y0 = u32sat(wmax(0.0f, __builtin_floorf((sy - r) / 16.0f)));
  x1 = u32sat(wmin((float)t16x - 1.0f, __builtin_floorf((sx + r) / 16.0f)));
  y1 = u32sat(wmin((float)t16y - 1.0f, __builtin_floorf((sy + r) / 16.0f)));
}

// Tile range of a splat at the grid's tile size (false: empty). Without
// maxCoverage truncation every size lists the splat in each tile holding a
// pixel that the reference blends it into:
//  16 px: the reference's range;
//  32 px: the 32-px tiles that hold a tile of the 16-px range (the
//         rasterizer keeps each 16-px quadrant to the splats whose 16-px
//         range holds it);
//   8 px: the 8-px tiles of the 16-px range, narrowed to floor((s -+ r) / 8)
//         when s -+ r and r are below 2^22 in magnitude. A pixel of the
//         16-px range outside the narrower range has its centre at least
//         1/2 beyond the rounded s -+ r, so at least 3/8 beyond the exact
//         one (the rounding error is at most 1/8 there), more than half the
//         float spacing at r (at most 1/8): the f32 difference px + 0.5 - s
//         exceeds r in magnitude and the reference's box test
//         (rasterize.ts:66-69) skips the splat at that pixel.
__host__ __device__ inline bool tile_range(float sx, float sy, float r, const TileGrid &g, uint32_t &x0,
                                           uint32_t &x1, uint32_t &y0, uint32_t &y1) {
  uint32_t a0, a1, b0, b1;
  tile_range16(sx, sy, r, g.t16x, g.t16y, a0, a1, b0, b1);
  if (a1 < a0 || b1 < b0) return false;
  if (g.tile == kTile) {
    x0 = a0; x1 = a1; y0 = b0; y1 = b1;
    return true;
  }
  if (g.tile == 32u) {
    x0 = a0 >> 1; x1 = a1 >> 1; y0 = b0 >> 1; y1 = b1 >> 1;
    return true;
  }
  x0 = 2u * a0;
  y0 = 2u * b0;
  x1 = 2u * a1 + 1u < g.tiles_x ? 2u * a1 + 1u : g.tiles_x - 1u;
  y1 = 2u * b1 + 1u < g.tiles_y ? 2u * b1 + 1u : g.tiles_y - 1u;
  const float ex = sx - r, fx = sx + r, ey = sy - r, fy = sy + r;
  const float lim = 4194304.0f;  // 2^22
  if (__builtin_fabsf(ex) < lim && __builtin_fabsf(fx) < lim && __builtin_fabsf(ey) < lim &&
      __builtin_fabsf(fy) < lim && r < lim) {
    const uint32_t nx0 = u32sat(wmax(0.0f, __builtin_floorf(ex / 8.0f)));
    const uint32_t ny0 = u32sat(wmax(0.0f, __builtin_floorf(ey / 8.0f)));
    const uint32_t nx1 = u32sat(__builtin_floorf(fx / 8.0f));
    const uint32_t ny1 = u32sat(__builtin_floorf(fy / 8.0f));
    x0 = x0 > nx0 ? x0 : nx0;
    y0 = y0 > ny0 ? y0 : ny0;
    x1 = x1 < nx1 ? x1 : nx1;
    y1 = y1 < ny1 ? y1 : ny1;
  }
  return x0 <= x1 && y0 <= y1;
}

// The rasterizer's strip test: can a pixel of a tile's columns [x0, x1] and
// a strip's rows [y0, y1] (pixel centres) pass the reference's box test
// |dx| <= r, |dy| <= r (rasterize.ts:66-69) with power >= tau, for a splat
// at (sx, sy) with conic (A, B, C)? "false" is a proof that none does.
//
// A pixel computes dx = fl(px - sx) in f32; rounding is monotone, so over
// the tile's columns dx lies in [fl(x0 - sx), fl(x1 - sx)], and it passes
// the box test only inside [-r, r]: the rectangle W of (dx, dy) is exact.
// power = -q/2 with q = A dx^2 + 2B dx dy + C dy^2 evaluated in f32 (three
// roundings per term), so a blended pixel has q(dx, dy) <= -2 tau + 3u mag,
// u = 2^-24, mag = A mx^2 + 2|B| mx my + C my^2 >= every term on W. With A,
// C > 0, q restricted to an edge of W is a convex parabola and q has no
// interior minimum but the centre, so the minimum of q on W is 0 (W holds
// the centre) or the least of the four edge minima. Each edge minimum is
// taken at its clamped vertex, computed from a rounded ratio (-B/C, -B/A):
// an error d in the vertex raises the value by at most 3 C d^2, a few u^2
// of mag. The candidates are evaluated in f32 (error below 4u mag). Their
// least value minus 1e-5 mag (and 1e-30 for underflow) therefore lies below
// the true minimum by more than the 3u mag a pixel needs: if it still
// exceeds -2 tau, no pixel of the rectangle is blended. NaN or infinite
// terms make the final comparison false (the strip is tested).
struct StripTest {
  float sy, r, A, B, C, kA, thr, dxl, dxh, yv0, yv1, mx;
  bool never, always;
};

// The strip-independent part, once per (splat, tile columns).
__host__ __device__ inline StripTest strip_test(float sx, float sy, float r, float A, float B, float C, float tau,
                                                float x0, float x1) {
  StripTest t;
  t.sy = sy; t.r = r; t.A = A; t.B = B; t.C = C;
  t.never = !(tau < __builtin_inff());
  t.always = !(A > 0.0f && C > 0.0f);
  t.dxl = __builtin_fmaxf(x0 - sx, -r);
  t.dxh = __builtin_fminf(x1 - sx, r);
  t.never = t.never || t.dxl > t.dxh;
  t.mx = __builtin_fmaxf(__builtin_fabsf(t.dxl), __builtin_fabsf(t.dxh));
  const float kC = -B / C;
  t.kA = -B / A;
  t.yv0 = kC * t.dxl;
  t.yv1 = kC * t.dxh;
  t.thr = -2.0f * tau;
  return t;
}

__host__ __device__ inline float strip_q(const StripTest &t, float x, float y) {
  return (t.A * x * x + 2.0f * t.B * x * y) + t.C * y * y;
}

__host__ __device__ inline bool strip_may_hit(const StripTest &t, float y0, float y1) {
  if (t.never) return false;
  if (t.always) return true;
  const float dyl = __builtin_fmaxf(y0 - t.sy, -t.r);
  const float dyh = __builtin_fminf(y1 - t.sy, t.r);
  if (dyl > dyh) return false;
  float qmin = 0.0f;
  if (!(t.dxl <= 0.0f && 0.0f <= t.dxh && dyl <= 0.0f && 0.0f <= dyh)) {
    const float ya = __builtin_fminf(__builtin_fmaxf(t.yv0, dyl), dyh);
    const float yb = __builtin_fminf(__builtin_fmaxf(t.yv1, dyl), dyh);
    const float xa = __builtin_fminf(__builtin_fmaxf(t.kA * dyl, t.dxl), t.dxh);
    const float xb = __builtin_fminf(__builtin_fmaxf(t.kA * dyh, t.dxl), t.dxh);
    qmin = __builtin_fminf(__builtin_fminf(strip_q(t, t.dxl, ya), strip_q(t, t.dxh, yb)),
                           __builtin_fminf(strip_q(t, xa, dyl), strip_q(t, xb, dyh)));
  }
  const float my = __builtin_fmaxf(__builtin_fabsf(dyl), __builtin_fabsf(dyh));
  const float mag = (t.A * t.mx * t.mx + 2.0f * __builtin_fabsf(t.B) * t.mx * my) + t.C * my * my;
  return !(qmin - 1e-5f * mag - 1e-30f > t.thr);
}

// floatToSortable (src/spatial/radix-sort.ts:20-24)
__host__ __device__ inline uint32_t sortable_key(float f) {
  uint32_t u = f32_bits(f);
  return (u & 0x80000000u) ? ~u : (u ^ 0x80000000u);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

}  // namespace dl
