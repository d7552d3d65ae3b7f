"""GPU checks of the PLY ingestion (dl_scene_create_from_ply, the device form
of loadPlySplats, src/io/ply.ts:120-239) against the CPU restatement
(oracle/ply.py): every SplatData array bit for bit on generated files
(all band layouts, shuffled and extra properties, odd data offsets, CRLF
and non-ASCII headers, degenerate quaternions), the same error messages for
malformed files, splatBounds (ply.ts:242-255), and renders of a PLY scene
equal to renders of the same SplatData uploaded."""
import numpy as np
import pytest

import oracle
from oracle import ply as P
from droplet_amd import synthetic as S
from droplet_amd.renderer import Renderer, Scene, load_ply_splats
from droplet_amd.types import SplatData

pytestmark = pytest.mark.gpu

BASE = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
        "rot_0", "rot_1", "rot_2", "rot_3"]


def _file(n, k_rest, seed, shuffle=False, extra=(), **kw):
    names = BASE + [f"f_rest_{k}" for k in range(k_rest)]
    rng = np.random.default_rng(seed)
    rows = rng.normal(0.0, 1.0, (n, len(names))).astype(np.float32)
    if n > 8:
        rows[0, 10:14] = 0.0
        rows[1, 10:14] = [np.inf, 1.0, 0.0, 0.0]
        rows[2, 10:14] = [np.nan, 1.0, 0.0, 0.0]
        rows[3, 10:14] = [1e-30, -2e-30, 3e-30, 0.0]
        rows[4, 10:14] = [3e37, 1e38, -2e38, 1e38]
        rows[5, 0:3] = [0.0, -0.0, np.nan]
    names = list(names) + list(extra)
    if extra:
        rows = np.concatenate([rows, rng.normal(0, 1, (n, len(extra))).astype(np.float32)], axis=1)
    if shuffle:
        perm = rng.permutation(len(names))
        names = [names[i] for i in perm]
        rows = rows[:, perm]
    return P.write_ply([("float", nm) for nm in names], rows, **kw)


def _same(a, b, what):
    """Equal bit for bit, except that any NaN matches any NaN."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=what)
    m = ~np.isnan(a)
    np.testing.assert_array_equal(a[m].view(np.uint32), b[m].view(np.uint32), err_msg=what)


def _compare(device, data, flip_y=True):
    ref = P.load_ply_splats(data, flip_y)
    sc = Scene.from_ply(device, data, flip_y)
    try:
        got = sc.read()
        assert sc.count == ref.count and sc.sh_bands == ref.sh_bands
        assert sc.sh_coeffs_per_channel == (ref.coeffs_per_channel if ref.sh_bands else 0)
        for f in ("positions", "rotations", "log_scales", "opacity_logits", "colors_dc"):
            _same(getattr(got, f), getattr(ref, f), f)
        if ref.sh_bands:
            _same(got.colors_sh, ref.colors_sh, "colors_sh")
        lo, hi = sc.bounds()
        rlo, rhi = P.splat_bounds(ref.positions)
        np.testing.assert_array_equal(np.array(lo), np.array(rlo))
        np.testing.assert_array_equal(np.array(hi), np.array(rhi))
        assert [np.copysign(1, v) for v in lo] == [np.copysign(1, v) for v in rlo]
        assert [np.copysign(1, v) for v in hi] == [np.copysign(1, v) for v in rhi]
    finally:
        sc.destroy()
    return ref


@pytest.mark.parametrize("k_rest", [0, 9, 12, 24, 45, 48])
def test_ply_arrays_match_oracle(device, k_rest):
    print(f"case: ply k_rest={k_rest}", flush=True)
    for flip in (True, False):
        _compare(device, _file(3000, k_rest, seed=k_rest, shuffle=True, extra=("nx", "ny", "nz")), flip)


def test_ply_header_variants(device):
    # odd data offsets, CRLF, a non-ASCII comment, a face element after the vertices
    d = _file(257, 9, seed=1, extra_header="comment x")
    _compare(device, d)
    d = _file(257, 9, seed=2, extra_header="comment café ☃ 𝄞")
    _compare(device, d)
    raw = _file(100, 3, seed=3, post_vertex="element face 0\nproperty list uchar int vertex_indices")
    cut = raw.index(b"end_header")
    _compare(device, raw[:cut].replace(b"\n", b"\r\n") + raw[cut:])
    # one row, and a large file (several staging workgroups and bound partials)
    _compare(device, _file(1, 45, seed=4))
    _compare(device, _file(300_001, 45, seed=5))


def test_ply_errors_match(device):
    rows = np.zeros((3, 14), np.float32)
    good = [("float", n) for n in BASE]
    cases = [b"ply\nformat binary_little_endian 1.0\n",
             b"plx\nend_header\n",
             P.write_ply(good, rows, fmt="ascii 1.0"),
             P.write_ply(good, rows, fmt="binary_big_endian 1.0"),
             P.write_ply(good, rows[:0]),
             P.write_ply(good + [("double", "extra")], np.zeros((3, 15), np.float32)),
             P.write_ply(good[:-1] + [("uchar", "rot_3")], rows),
             P.write_ply(good, rows)[:-1],
             P.write_ply(good[1:], rows[:, 1:]),
             b"ply\nformat binary_little_endian 1.0\nelement vertex abc\nend_header\n"]
    for data in cases:
        with pytest.raises(P.PlyError) as want:
            P.load_ply_splats(data)
        with pytest.raises(Exception) as got:
            Scene.from_ply(device, data)
        assert str(want.value) in str(got.value), (str(want.value), str(got.value))


def test_ply_scene_renders_like_uploaded_splatdata(device):
    """A bands-1 file with 4 coefficients per channel (k_rest = 12): the
    renderer reads colorsSH with the stride of shBands, as the reference's
    packer does, for the PLY scene and for the same SplatData uploaded."""
    rng = np.random.default_rng(6)
    src = S.gaussian_cloud(20000, seed=6, sh_bands=1)
    n = src.count
    names = BASE + [f"f_rest_{k}" for k in range(12)]
    rows = np.zeros((n, len(names)), np.float32)
    pos = src.positions.reshape(n, 3)
    rows[:, 0:3] = pos * np.array([1, -1, 1], np.float32)
    rows[:, 3:6] = src.log_scales.reshape(n, 3)
    rows[:, 6:9] = src.colors_dc.reshape(n, 3) / np.float32(P.SH_C0)
    rows[:, 9] = src.opacity_logits
    rows[:, 10:14] = src.rotations.reshape(n, 4) * np.array([1, -1, 1, -1], np.float32)
    rows[:, 14:] = rng.normal(0, 0.1, (n, 12)).astype(np.float32)
    data = P.write_ply([("float", nm) for nm in names], rows)
    ref = P.load_ply_splats(data)
    sd = SplatData(ref.count, ref.positions, ref.rotations, ref.log_scales, ref.opacity_logits, ref.colors_dc,
                   ref.colors_sh, ref.sh_bands)
    cam = S.default_camera(320, 240)
    r = Renderer(device)
    a = Scene.from_ply(device, data)
    b = Scene(device, sd)
    try:
        img_a = r.render_scene(a, cam).copy()
        img_b = r.render_scene(b, cam).copy()
    finally:
        a.destroy()
        b.destroy()
        r.destroy()
    np.testing.assert_array_equal(img_a, img_b)
    img_o = oracle.render(sd, cam, taps=False)["rgba"].reshape(img_a.shape)
    np.testing.assert_array_equal(img_a, img_o)


def test_load_ply_splats_from_path(device, tmp_path):
    data = _file(500, 24, seed=7)
    p = tmp_path / "scene.ply"
    p.write_bytes(data)
    got = load_ply_splats(str(p), device=device)
    ref = P.load_ply_splats(data)
    _same(got.colors_sh, ref.colors_sh, "colors_sh")
    _same(got.rotations, ref.rotations, "rotations")
