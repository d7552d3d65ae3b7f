"""Pins the CPU restatement of the PLY loader (oracle/ply.py, the checker of
the device PLY ingestion) against what the reference's loader fixes:
the error strings of src/io/ply.ts, the band thresholds of ply.ts:170-177,
property-order independence, the Y flip, unit quaternions, the SH_C0
scaling, Math.hypot against a library hypot, and splatBounds against
numpy's NaN-skipping min/max."""
import math

import numpy as np
import pytest

from oracle import ply as P

BASE = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
        "rot_0", "rot_1", "rot_2", "rot_3"]


def _props(names):
    return [("float", n) for n in names]


def _rows(n, names, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 1.0, (n, len(names))).astype(np.float32)


def test_error_strings():
    good = _props(BASE)
    rows = _rows(3, BASE, 0)
    with pytest.raises(P.PlyError, match="end_header not found in first 64KB"):
        P.parse_header(b"ply\nformat binary_little_endian 1.0\n")
    with pytest.raises(P.PlyError, match="missing magic"):
        P.parse_header(b"plx\nend_header\n")
    with pytest.raises(P.PlyError, match="ASCII PLY not supported"):
        P.load_ply_splats(P.write_ply(good, rows, fmt="ascii 1.0"))
    with pytest.raises(P.PlyError, match="only little-endian supported"):
        P.load_ply_splats(P.write_ply(good, rows, fmt="binary_big_endian 1.0"))
    with pytest.raises(P.PlyError, match="no vertices"):
        P.load_ply_splats(P.write_ply(good, rows[:0]))
    with pytest.raises(P.PlyError, match='unsupported property type "double"'):
        P.load_ply_splats(P.write_ply(good + [("double", "extra")], np.zeros((3, 15), np.float32)))
    with pytest.raises(P.PlyError, match="mixed property types not yet supported"):
        P.load_ply_splats(P.write_ply(good[:-1] + [("uchar", "rot_3")], rows))
    with pytest.raises(P.PlyError, match=r"data section truncated \(need 168 bytes, have 167\)"):
        P.load_ply_splats(P.write_ply(good, rows)[:-1])
    with pytest.raises(P.PlyError, match='missing property "opacity"'):
        names = [n for n in BASE if n != "opacity"]
        P.load_ply_splats(P.write_ply(_props(names), _rows(3, names, 1)))


def test_band_detection_and_sh_layout():
    for k_rest, bands in [(0, 0), (2, 0), (3, 0), (9, 1), (10, 1), (12, 1), (24, 2), (44, 2), (45, 3), (48, 3)]:
        names = BASE + [f"f_rest_{k}" for k in range(k_rest)]
        rows = _rows(5, names, k_rest)
        d = P.load_ply_splats(P.write_ply(_props(names), rows))
        assert d.sh_bands == bands and d.coeffs_per_channel == k_rest // 3
        if bands:
            sh = d.colors_sh.reshape(5, -1)
            np.testing.assert_array_equal(sh, rows[:, 14:14 + 3 * (k_rest // 3)])
        else:
            assert d.colors_sh is None


def test_property_order_and_extra_properties():
    names = BASE + [f"f_rest_{k}" for k in range(9)]
    rows = _rows(200, names, 7)
    a = P.load_ply_splats(P.write_ply(_props(names), rows))
    perm = np.random.default_rng(8).permutation(len(names))
    extra = ["nx", "ny", "nz"]
    names2 = [names[i] for i in perm] + extra
    rows2 = np.concatenate([rows[:, perm], np.ones((200, 3), np.float32)], axis=1)
    b = P.load_ply_splats(P.write_ply(_props(names2), rows2, pre_vertex="comment made by a test",
                                      post_vertex="element face 0\nproperty list uchar int vertex_indices"))
    for f in ("positions", "rotations", "log_scales", "opacity_logits", "colors_dc", "colors_sh"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))


def test_values_flip_and_normalisation():
    names = list(BASE)
    rows = _rows(1000, names, 9)
    rows[0, 10:14] = 0.0                     # zero quaternion: `|| 1`
    rows[1, 10:14] = [3.0, 0.0, 4.0, 0.0]
    data = P.write_ply(_props(names), rows)
    up = P.load_ply_splats(data, flip_y=False)
    dn = P.load_ply_splats(data, flip_y=True)
    pos = rows[:, 0:3]
    np.testing.assert_array_equal(up.positions.reshape(-1, 3), pos)
    np.testing.assert_array_equal(dn.positions.reshape(-1, 3), pos * np.array([1, -1, 1], np.float32))
    np.testing.assert_array_equal(up.log_scales.reshape(-1, 3), rows[:, 3:6])
    np.testing.assert_array_equal(up.opacity_logits, rows[:, 9])
    # f_dc * SH_C0 with SH_C0 = 1 / (2 sqrt(pi)), rounded once to f32
    sh_c0 = 1.0 / (2.0 * math.sqrt(math.pi))
    assert abs(sh_c0 - P.SH_C0) < 1e-16
    dc = np.array([np.float32(sh_c0 * float(v)) for v in rows[:, 6:9].reshape(-1)], np.float32)
    np.testing.assert_array_equal(up.colors_dc, dc)
    q = up.rotations.reshape(-1, 4)
    np.testing.assert_array_equal(q[0], np.zeros(4, np.float32))
    np.testing.assert_array_equal(q[1], np.array([0.6, 0.0, 0.8, 0.0], np.float32))
    norms = np.linalg.norm(q[2:].astype(np.float64), axis=1)
    assert np.abs(norms - 1.0).max() < 1e-6
    # the flip negates x and z of the quaternion and nothing else
    qd = dn.rotations.reshape(-1, 4)
    np.testing.assert_array_equal(qd, q * np.array([1, -1, 1, -1], np.float32))
    # each component is the f32 rounding of component / |q| (library hypot)
    ref = np.array([[np.float32(float(c) / math.hypot(*map(float, r))) for c in r] for r in rows[2:60, 10:14]])
    assert (np.abs(q[2:60].view(np.int32) - ref.view(np.int32)) <= 1).all()


def test_js_hypot_against_library():
    rng = np.random.default_rng(10)
    cols = [rng.normal(0, 1, 20000) * 10.0 ** rng.integers(-30, 30, 20000) for _ in range(4)]
    h = P.js_hypot(*cols)
    lib = np.array([math.hypot(*(float(c[i]) for c in cols)) for i in range(20000)])
    assert (np.abs(h - lib) <= 2 * np.spacing(lib)).all()
    z = np.zeros(3)
    np.testing.assert_array_equal(P.js_hypot(z, z, z, z), z)
    assert np.isinf(P.js_hypot(np.array([np.inf]), np.array([np.nan]), z[:1], z[:1]))[0]
    assert np.isnan(P.js_hypot(np.array([1.0]), np.array([np.nan]), z[:1], z[:1]))[0]


def test_header_quirks():
    names = list(BASE)
    rows = _rows(4, names, 11)
    plain = P.write_ply(_props(names), rows)
    # CRLF line ends and stray spaces are trimmed; `element vertex 4 trailing` parses as 4
    cut = plain.index(b"end_header")
    crlf = plain[:cut].replace(b"\n", b"\r\n").replace(b"element vertex 4", b"element vertex   4 trailing") + \
        plain[cut:]
    assert P.load_ply_splats(crlf).count == 4
    # the data offset counts UTF-16 units of the decoded header: a 2-byte
    # UTF-8 character in a comment moves it one byte earlier than the byte offset
    hdr = P.parse_header(P.write_ply(_props(names), rows, extra_header="comment café"))
    raw = P.write_ply(_props(names), rows, extra_header="comment café")
    assert hdr.data_offset == raw.index(b"end_header\n") + 11 - 1


def test_splat_bounds():
    rng = np.random.default_rng(12)
    p = rng.normal(0, 3, (500, 3)).astype(np.float32)
    p[7, 1] = np.nan
    lo, hi = P.splat_bounds(p.reshape(-1))
    np.testing.assert_array_equal(lo, np.nanmin(p.astype(np.float64), axis=0))
    np.testing.assert_array_equal(hi, np.nanmax(p.astype(np.float64), axis=0))
    lo, hi = P.splat_bounds(np.zeros(0, np.float32))
    assert lo == [math.inf] * 3 and hi == [-math.inf] * 3
    # equal values: the first in index order stays (the sign of a zero)
    z = np.array([[0.0, -0.0, 1.0], [-0.0, 0.0, 1.0]], np.float32)
    lo, hi = P.splat_bounds(z.reshape(-1))
    assert [math.copysign(1, v) for v in lo[:2]] == [1, -1] and [math.copysign(1, v) for v in hi[:2]] == [1, -1]
