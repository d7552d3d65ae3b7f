"""The voxelizer ends a voxel's Gaussian sum at voxel_solid_stop(cutoff)
(droplet_amd/csrc/dl_common.h) instead of the reference's sigma >= 7
(src/voxel/shaders/voxelize.ts:111-130). That leaves every mask bit unchanged
if the reference's solid test (:132-134), 1 - exp(-sigma) >= cutoff in f32,
holds for every float sigma from the stop value up to 8, the largest sum the
reference can reach (it stops at 7 and adds terms of at most 1). This checks
that property for every float in [stop, 8] with dl_expf, for the workload's
cutoff and cutoffs near both ends of (0, 1); that the stop lies within
0.05% of the exact threshold -ln(1 - c), so the shortcut is not wasted;
and that the stop is 7 for cutoffs outside (0, 1). The checker is built
with hipcc (host code only) from the header itself."""
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
#include <cstring>
#include "dl_common.h"
using namespace dl;
static uint32_t u(float x) { uint32_t v; std::memcpy(&v, &x, 4); return v; }
static float f(uint32_t b) { float x; std::memcpy(&x, &b, 4); return x; }
int main() {
  unsigned long long bad = 0, n = 0;
  const float cuts[] = {0.1f, 0.5f, 0.01f, 1e-6f, 1e-30f, 0.9f, 0.99f, 0.999f, 0.3f, 0.75f};
  for (float c : cuts) {
    const float stop = voxel_solid_stop(c);
    const double exact = -std::log(1.0 - (double)c);
    if (stop < 7.0f && !(stop >= exact && stop <= exact * 1.0005 + 1e-6)) {
      printf("cutoff %a: stop %a not within 0.05%% above %a\n", c, stop, exact);
      bad++;
    }
    unsigned long long nc = 0;
    for (uint32_t b = u(stop); b <= u(8.0f); b++) {
      const float sigma = f(b);
      const float opacity = 1.0f - dl_expf(-sigma);
      if (!(opacity >= c)) {
        if (bad < 10) printf("cutoff %a: sigma %a gives opacity %a < cutoff\n", c, sigma, opacity);
        bad++;
      }
      nc++;
    }
    n += nc;
    printf("cutoff %g stop %.9g (exact %.9g) checked %llu\n", c, stop, exact, nc);
  }
  const float outside[] = {0.0f, -0.5f, 1.0f, 1.5f, NAN, INFINITY};
  for (float c : outside)
    if (voxel_solid_stop(c) != 7.0f) { printf("cutoff %a: stop %a, not 7\n", c, voxel_solid_stop(c)); bad++; }
  printf("checked %llu failures %llu\n", n, bad);
  return bad ? 1 : 0;
}
"""


def test_voxel_solid_stop_keeps_every_mask_bit():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not available")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "stopcheck.cpp"), os.path.join(d, "stopcheck")
        open(src, "w").write(PROG)
        subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                               "--cuda-host-only", "-I", os.path.join(ROOT, "droplet_amd", "csrc"), src, "-o", exe],
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "failures 0" in out.stdout
