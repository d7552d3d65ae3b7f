"""CPU checks of the boundary: the C-ABI library loads without a GPU, exports
every entry point include/droplet_hip.h declares, the ctypes signature table
covers them all, and the Python mirror validates its inputs like the
reference's typed arrays do. No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from droplet_amd import _lib as L
from droplet_amd import synthetic as S
from droplet_amd.types import RenderCamera, SplatData, splat_input_stride

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "droplet_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dl_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_signature_table_matches_header():
    assert sorted(n for n, _, _ in L.SIGNATURES) == _declared()
    L.load()  # binds every signature


def test_status_codes_and_error_string():
    lib = L.load()
    # a null out-pointer is rejected before any device call
    rc = lib.dl_context_create(None, None)
    assert rc == L.DL_ERR_INVALID
    assert b"null" in lib.dl_last_error()


def test_splatdata_validation():
    with pytest.raises(ValueError):
        SplatData(2, np.zeros(5), np.zeros(8), np.zeros(6), np.zeros(2), np.zeros(6))
    with pytest.raises(ValueError):
        SplatData(1, np.zeros(3), np.zeros(4), np.zeros(3), np.zeros(1), np.zeros(3), None, 2)
    with pytest.raises(ValueError):
        RenderCamera(10, 10, np.zeros(16), 1.0, 0.1, 10.0, "fisheye")
    assert splat_input_stride(3) == 59 and splat_input_stride(0) == 14


def test_struct_layouts_match_header():
    # plain-C layouts that bun:ffi / ctypes bind to
    assert ctypes.sizeof(L.DlSplats) == 8 + 6 * 8 + 8
    assert ctypes.sizeof(L.DlCamera) == 8 + 64 + 24 + 8
    assert ctypes.sizeof(L.DlBackground) == 16
    assert ctypes.sizeof(L.DlRendererOpts) == 36
    assert ctypes.sizeof(L.DlView) == 8 + 104 + 16
    assert ctypes.sizeof(L.DlBatchStats) == 32
    assert ctypes.sizeof(L.DlVoxelOpts) == 8 + 8 + 8 + 48


def test_generators_are_deterministic():
    a = S.gaussian_cloud(1000, seed=9, sh_bands=3)
    b = S.gaussian_cloud(1000, seed=9, sh_bands=3)
    for f in ("positions", "rotations", "log_scales", "opacity_logits", "colors_dc", "colors_sh"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    q = a.rotations.reshape(-1, 4).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    assert np.abs(a.positions).max() <= 2.0


def test_voxel_device_view_layout():
    # masks, num_blocks, nb[3] (+4 pad), bounds_min[3], voxel_resolution,
    # three list pointers, count, device (+4 pad)
    assert ctypes.sizeof(L.DlVoxelDeviceView) == 8 + 8 + 16 + 24 + 8 + 24 + 8 + 8
    assert L.DlVoxelDeviceView.bounds_min.offset == 32
    assert L.DlVoxelDeviceView.count.offset == 88
