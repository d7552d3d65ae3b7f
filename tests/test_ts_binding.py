"""The TypeScript drop-in (ts/droplet_hip.ts) against the C ABI it binds
(include/droplet_hip.h), on the CPU: bun is not in this image, so the
binding cannot run here, but every symbol it declares must exist in the
header with the same arguments (count and kind) and return kind, every
function of the header must be bound, and the struct sizes and field
offsets it writes must be the C compiler's (a probe compiled with gcc
against the header), which the Python ctypes mirror must match too."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from droplet_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "droplet_hip.h")
TS = os.path.join(ROOT, "ts", "droplet_hip.ts")


def _strip_c_comments(s):
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def header_functions():
    src = _strip_c_comments(open(HEADER).read())
    src = "\n".join(line for line in src.splitlines() if not line.lstrip().startswith("#"))
    out = {}
    for m in re.finditer(r"((?:const\s+)?[A-Za-z_]\w*\s*\**)\s*\b(dl_\w+)\s*\(([^;{}]*?)\)\s*;", src):
        ret, name, params = " ".join(m.group(1).split()), m.group(2), " ".join(m.group(3).split())
        if name.startswith("dl_") and not ret.startswith("typedef"):
            ps = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
            out[name] = (ret, ps)
    return out


def ts_symbols():
    src = open(TS).read()
    block = src[src.index("export const symbols = {"):]
    block = block[:block.index("} as const;")]
    out = {}
    for m in re.finditer(r"(dl_\w+): \{ args: \[([^\]]*)\], returns: ([\w.]+) \}", block):
        args = [a.strip() for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = (args, m.group(3))
    return out


def _kind_of_param(p):
    p = p.replace("const ", "")
    if "*" in p or "[" in p:
        return "P"
    t = p.split()[0]
    return {"uint32_t": "U32", "int": "I32", "uint64_t": "U64", "size_t": "U64", "double": "FFIType.f64"}[t]


def _kind_of_return(r):
    r = r.replace("const ", "")
    if "char *" in r or "char*" in r:
        return "FFIType.cstring"
    return {"int": "I32", "void": "VOID", "uint64_t": "U64", "float": "F32"}[r.strip()]


def test_every_ts_symbol_matches_the_header():
    hdr, ts = header_functions(), ts_symbols()
    assert len(hdr) > 50 and len(ts) > 50
    missing = sorted(set(hdr) - set(ts))
    extra = sorted(set(ts) - set(hdr))
    assert not missing, f"header functions without a TS binding: {missing}"
    assert not extra, f"TS symbols the header does not declare: {extra}"
    for name, (ret, params) in hdr.items():
        args, rk = ts[name]
        assert len(args) == len(params), f"{name}: {len(args)} TS args, {len(params)} in the header"
        assert args == [_kind_of_param(p) for p in params], (name, args, params)
        assert rk == _kind_of_return(ret), (name, rk, ret)


def test_every_header_function_is_exported():
    lib = C.CDLL(L.LIB_PATH) if os.path.exists(L.LIB_PATH) else None
    if lib is None:
        pytest.skip("libdroplet_hip.so is not built")
    for name in header_functions():
        assert hasattr(lib, name), name


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "droplet_hip.h"
#define S(t) printf(#t " size %zu\n", sizeof(t));
#define O(t, f) printf(#t " " #f " %zu\n", offsetof(t, f));
int main(void) {
  S(dl_splats) O(dl_splats, positions) O(dl_splats, colors_sh) O(dl_splats, sh_bands)
  S(dl_camera) O(dl_camera, view_matrix) O(dl_camera, fov_y) O(dl_camera, near_plane) O(dl_camera, far_plane)
  O(dl_camera, projection)
  S(dl_background) S(dl_renderer_opts) O(dl_renderer_opts, tile_size) O(dl_renderer_opts, raster_group)
  S(dl_device_opts)
  S(dl_voxel_opts) O(dl_voxel_opts, opacity_cutoff) O(dl_voxel_opts, has_bounds) O(dl_voxel_opts, bounds_min)
  O(dl_voxel_opts, bounds_max)
  S(dl_grads) S(dl_view) O(dl_view, camera) O(dl_view, background)
  S(dl_tune_result) O(dl_tune_result, chosen_ms) O(dl_tune_result, default_ms) O(dl_tune_result, candidates)
  S(dl_render_stats) O(dl_render_stats, num_pairs) O(dl_render_stats, stage_ms) O(dl_render_stats, total_ms)
  O(dl_render_stats, pixel_iters) O(dl_render_stats, list_entries) O(dl_render_stats, depth_key_bits)
  O(dl_render_stats, tile_size) O(dl_render_stats, raster_group)
  S(dl_batch_stats) S(dl_voxel_stats) S(dl_voxel_device_view)
  return 0;
}
"""


@pytest.fixture(scope="module")
def c_layout():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write(PROBE)
        subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe])
        out = subprocess.check_output([exe], text=True)
    lay = {}
    for line in out.splitlines():
        parts = line.split()
        lay[" ".join(parts[:-1])] = int(parts[-1])
    return lay


def test_ts_layouts_are_the_c_layouts(c_layout):
    src = open(TS).read()
    block = src[src.index("export const LAYOUT = {"):]
    block = block[:block.index("} as const;")]
    ts = {m.group(1): int(m.group(2)) for m in re.finditer(r"(dl_\w+): (\d+),", block)}
    assert len(ts) >= 12
    for name, size in ts.items():
        assert c_layout[f"{name} size"] == size, name
    # the offsets the TS packers write
    want = {"dl_splats positions": 8, "dl_splats colors_sh": 48, "dl_splats sh_bands": 56,
            "dl_camera view_matrix": 8, "dl_camera fov_y": 72, "dl_camera near_plane": 80,
            "dl_camera far_plane": 88, "dl_camera projection": 96,
            "dl_voxel_opts opacity_cutoff": 8, "dl_voxel_opts has_bounds": 16, "dl_voxel_opts bounds_min": 24,
            "dl_voxel_opts bounds_max": 48, "dl_view camera": 8, "dl_view background": 112,
            "dl_tune_result chosen_ms": 36, "dl_tune_result default_ms": 40, "dl_tune_result candidates": 44,
            "dl_renderer_opts tile_size": 28, "dl_renderer_opts raster_group": 32,
            "dl_render_stats tile_size": 124, "dl_render_stats raster_group": 128,
            "dl_render_stats num_pairs": 24, "dl_render_stats stage_ms": 32, "dl_render_stats total_ms": 64,
            "dl_render_stats pixel_iters": 72, "dl_render_stats list_entries": 88}
    for k, v in want.items():
        assert c_layout[k] == v, k


def test_python_mirror_layouts_are_the_c_layouts(c_layout):
    pairs = {"dl_splats": L.DlSplats, "dl_camera": L.DlCamera, "dl_background": L.DlBackground,
             "dl_renderer_opts": L.DlRendererOpts, "dl_device_opts": L.DlDeviceOpts, "dl_voxel_opts": L.DlVoxelOpts,
             "dl_grads": L.DlGrads, "dl_view": L.DlView, "dl_tune_result": L.DlTuneResult,
             "dl_render_stats": L.DlRenderStats, "dl_batch_stats": L.DlBatchStats, "dl_voxel_stats": L.DlVoxelStats,
             "dl_voxel_device_view": L.DlVoxelDeviceView}
    for name, cls in pairs.items():
        assert C.sizeof(cls) == c_layout[f"{name} size"], name
    assert L.DlRenderStats.depth_key_bits.offset == c_layout["dl_render_stats depth_key_bits"]
    assert L.DlView.background.offset == c_layout["dl_view background"]
