"""The rasterizer's strip test (strip_test / strip_may_hit in
droplet_amd/csrc/dl_common.h) lets a wave skip a splat on a 16x4 strip of
pixels when no pixel of the strip can blend it. Skipping is only exact if
the test never answers "no" for a strip holding a pixel that the
reference's loop body (rasterize.ts:63-78: the r-box test, then power >=
tau with the power of rasterize.ts:71 in f32) would blend. This checks that
on the CPU, by brute force over every pixel of every rejected strip:

- random splats shaped like the projection's output (the AA-widened 2D
  covariance, its conic, radius = ceil(3 sqrt(lambda_max)), the skip
  threshold tau of an opacity), at random offsets around a tile, at small
  and at large screen coordinates;
- adversarial thresholds: tau set to the exact f32 power of one pixel of
  the strip (that pixel is then blended), and r set to the exact |dx| or
  |dy| of a pixel (that pixel sits on the box edge);
- degenerate and special inputs (r = 0, tiny and huge conics, NaN and
  infinite terms), where the test must not reject a strip that a pixel
  passes. tau itself is +inf (alpha <= 0: never blended) or finite, as
  project_splat computes it, so a NaN tau is not among them.

It also reports how often the test answers "yes" for a strip that no pixel
passes (its looseness). The checker is built with hipcc (host code only)
from the header itself."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROG = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include "dl_common.h"
using namespace dl;

struct Splat { float sx, sy, r, A, B, C, tau; };

// the reference's loop body up to the blend decision (rasterize.ts:63-78 with
// this build's skip threshold tau, which only drops alphas below 1/255)
static bool pixel_blends(const Splat &s, float pxf, float pyf) {
  const float dx = pxf - s.sx, dy = pyf - s.sy;
  if (!(s.r > 0.0f) || std::fabs(dx) > s.r || std::fabs(dy) > s.r) return false;
  const float power = (-0.5f) * std::fmaf(s.C * dy, dy, std::fmaf((2.0f * s.B) * dx, dy, (s.A * dx) * dx));
  return !(power > 0.0f || power < s.tau);
}

static unsigned long long n_strips = 0, n_rejected = 0, n_loose = 0, n_bad = 0;

// one 16x4 strip of the tile at (ox, oy), strip k
static void check(const Splat &s, uint32_t ox, uint32_t oy, int k) {
  const float x0 = (float)ox + 0.5f, x1 = x0 + 15.0f;
  const float y0 = (float)(oy + 4 * k) + 0.5f;
  const bool may = strip_may_hit(strip_test(s.sx, s.sy, s.r, s.A, s.B, s.C, s.tau, x0, x1), y0, y0 + 3.0f);
  bool any = false;
  for (int yy = 0; yy < 4 && !any; yy++)
    for (int xx = 0; xx < 16 && !any; xx++)
      any = pixel_blends(s, (float)(ox + xx) + 0.5f, (float)(oy + 4 * k + yy) + 0.5f);
  n_strips++;
  if (!may) n_rejected++;
  if (may && !any) n_loose++;
  if (!may && any) {
    if (n_bad < 10)
      printf("UNSOUND sx=%a sy=%a r=%a A=%a B=%a C=%a tau=%a tile=(%u,%u) strip %d\n", s.sx, s.sy, s.r, s.A, s.B,
             s.C, s.tau, ox, oy, k);
    n_bad++;
  }
}

// a splat shaped like project_splat's record: covariance sigma-scaled and
// rotated, widened by 0.3 on the diagonal, its conic, radius and tau
static Splat make_splat(std::mt19937_64 &g, float sx, float sy) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double ls1 = -2.0 + 7.0 * U(g), ls2 = -2.0 + 7.0 * U(g);  // log sigma in px: 0.14 .. 150
  const float s1 = (float)std::exp(ls1), s2 = (float)std::exp(ls2);
  const float th = (float)(6.283185307179586 * U(g));
  const float c = std::cos(th), sn = std::sin(th);
  float cov00 = c * c * s1 * s1 + sn * sn * s2 * s2;
  const float cov01 = c * sn * (s1 * s1 - s2 * s2);
  float cov11 = sn * sn * s1 * s1 + c * c * s2 * s2;
  cov00 = cov00 + 0.3f;
  cov11 = cov11 + 0.3f;
  const float det = cov00 * cov11 - cov01 * cov01;
  Splat s;
  s.sx = sx; s.sy = sy;
  const float inv = 1.0f / det;
  s.A = cov11 * inv; s.B = (-cov01) * inv; s.C = cov00 * inv;
  const float mid = 0.5f * (cov00 + cov11);
  const float l1 = mid + std::sqrt(std::fmax(0.1f, mid * mid - det));
  s.r = std::ceil(3.0f * std::sqrt(l1));
  const double alpha = std::pow(10.0, -2.5 + 2.5 * U(g));  // 0.003 .. 1
  const double target = (double)kGaussFloor + ((double)kMinAlpha / alpha) * (1.0 - 1e-6);
  const double t = std::log(target) - 1e-5;
  s.tau = (float)(t - std::fabs(t) * 1e-6 - 1e-30);
  return s;
}

int main() {
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  // 1. random splats around random tiles, small and large coordinates
  for (int it = 0; it < 400000; it++) {
    const uint32_t scale = (it & 3) == 3 ? 1u << 20 : 4096u;
    const uint32_t ox = 16u * (uint32_t)(U(g) * (scale / 16)), oy = 16u * (uint32_t)(U(g) * (scale / 16));
    Splat s = make_splat(g, 0.0f, 0.0f);
    const float reach = s.r + 16.0f;
    s.sx = (float)((double)ox + 8.0 + (2.0 * U(g) - 1.0) * reach);
    s.sy = (float)((double)oy + 8.0 + (2.0 * U(g) - 1.0) * reach);
    if ((it & 7) == 5) s.sx = (float)ox + (float)(int)(U(g) * 16) + 0.5f;  // centre on a pixel centre
    for (int k = 0; k < 4; k++) check(s, ox, oy, k);
  }
  // 2. adversarial tau: exactly one pixel's f32 power (that pixel blends)
  for (int it = 0; it < 300000; it++) {
    const uint32_t ox = 16u * (uint32_t)(U(g) * 256), oy = 16u * (uint32_t)(U(g) * 256);
    Splat s = make_splat(g, 0.0f, 0.0f);
    const float reach = s.r + 8.0f;
    s.sx = (float)((double)ox + 8.0 + (2.0 * U(g) - 1.0) * reach);
    s.sy = (float)((double)oy + 8.0 + (2.0 * U(g) - 1.0) * reach);
    const int k = (int)(U(g) * 4), xx = (int)(U(g) * 16), yy = (int)(U(g) * 4);
    const float pxf = (float)(ox + xx) + 0.5f, pyf = (float)(oy + 4 * k + yy) + 0.5f;
    const float dx = pxf - s.sx, dy = pyf - s.sy;
    if ((it & 3) == 1) s.r = std::fmax(std::fabs(dx), std::fabs(dy));  // the pixel on the box edge
    const float power = (-0.5f) * std::fmaf(s.C * dy, dy, std::fmaf((2.0f * s.B) * dx, dy, (s.A * dx) * dx));
    if (!(power <= 0.0f)) continue;
    s.tau = power;
    if ((it & 3) == 2) s.tau = std::nextafter(power, INFINITY);
    check(s, ox, oy, k);
  }
  // 3. special inputs
  const float specials[] = {0.0f, -0.0f, 1e-38f, 1e-30f, 1e30f, INFINITY, -INFINITY, NAN};
  for (float v : specials) {
    for (int which = 0; which < 7; which++) {
      Splat s = make_splat(g, 100.3f, 50.7f);
      float *f[7] = {&s.sx, &s.sy, &s.r, &s.A, &s.B, &s.C, &s.tau};
      // tau is +inf or finite by construction (project_splat: the log of a
      // finite positive target); a NaN tau is outside the contract
      if (which == 6 && std::isnan(v)) continue;
      *f[which] = v;
      for (int k = 0; k < 4; k++) check(s, 96, 48, k);
    }
  }
  printf("strips %llu rejected %llu loose %llu unsound %llu\n", n_strips, n_rejected, n_loose, n_bad);
  return n_bad ? 1 : 0;
}
"""


def test_strip_test_never_rejects_a_blended_pixel():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not available")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "stripcheck.cpp"), os.path.join(d, "stripcheck")
        open(src, "w").write(PROG)
        subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                               "--cuda-host-only", "-I", os.path.join(ROOT, "droplet_amd", "csrc"), src, "-o", exe],
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    fields = dict(zip(out.stdout.split()[0::2], out.stdout.split()[1::2]))
    assert int(fields["unsound"]) == 0
    # the test must reject a real share of the strips, or it skips nothing
    assert int(fields["rejected"]) > int(fields["strips"]) // 10
