"""The voxelizer evaluates e^x of x = -dd/2 <= 0 as
dl_expf_nonpos(x < -104 ? -104 : x) instead of the exp specification
dl_expf (droplet_amd/csrc/dl_common.h); the rasterizer uses dl_expf_nonpos on
[tau, 0]. This checks on the CPU, for every float in [-104, -0.5] and for
-inf, NaN and values below -104, that the two agree bit for bit (NaN
results: both NaN). Above -0.5 both take k = 0, where dl_expf_nonpos scales
by 2^0 and dl_expf by 2^0 * 2^0, so both return the same polynomial value;
a sample of that range is checked too. The checker is built with hipcc
(host code only) from the header itself."""
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
static float f(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }
static uint32_t u(float x) { uint32_t v; std::memcpy(&v, &x, 4); return v; }
static float voxel_exp(float x) { return dl_expf_nonpos(x < -104.0f ? -104.0f : x); }
static int same(float a, float b) { return (std::isnan(a) && std::isnan(b)) || u(a) == u(b); }
int main() {
  unsigned long long n = 0, bad = 0;
  // negative floats from -0.5 (0xBF000000) to -104 (0xC2D00000): every bit pattern
  for (uint32_t b = 0xBF000000u; b <= 0xC2D00000u; b++) {
    const float x = f(b);
    if (!same(dl_expf(x), voxel_exp(x))) { if (bad < 5) printf("mismatch %a\n", x); bad++; }
    n++;
  }
  // (-0.5, 0]: every 97th pattern, and the boundaries
  for (uint32_t b = 0x80000000u; b < 0xBF000000u; b += 97u) {
    const float x = f(b);
    if (!same(dl_expf(x), voxel_exp(x))) { if (bad < 5) printf("mismatch %a\n", x); bad++; }
    n++;
  }
  const float special[] = {0.0f, -0.0f, -1e-45f, -104.0f, -104.5f, -200.0f, -1e30f, -INFINITY, NAN, -NAN, -87.0f, -87.4f, -103.99f};
  for (float x : special) {
    if (!same(dl_expf(x), voxel_exp(x))) { printf("mismatch special %a\n", x); bad++; }
    n++;
  }
  printf("checked %llu mismatches %llu\n", n, bad);
  return bad ? 1 : 0;
}
"""


def test_voxel_exp_equals_the_exp_specification():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not available")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "expcheck.cpp"), os.path.join(d, "expcheck")
        open(src, "w").write(PROG)
        subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                               "--cuda-host-only", "-I", os.path.join(ROOT, "droplet_amd", "csrc"), src, "-o", exe],
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "mismatches 0" in out.stdout
