"""Pins the oracle's primitives against libm and closed forms (CPU only).

* orc_expf: <= 1 ulp from the correctly rounded e^x over the full f32 range
  the path uses, and exactly monotone there.
* float_to_sortable: order-preserving (src/spatial/radix-sort.ts:20-24).
* u32sat: WGSL u32(f32) round-toward-zero + clamp.
* focal: H / (2 tan(fovY/2)) (src/render/camera.ts:34).
* Morton encode/decode round trip (src/voxel/morton.ts:11-22).
"""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S


def _ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def test_expf_within_one_ulp_of_correctly_rounded():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        np.linspace(-103.0, 88.7, 2_000_001, dtype=np.float64),
        rng.uniform(-5.0, 0.0, 2_000_000),          # rasterize range of power
        rng.uniform(-8.0, 8.0, 1_000_000),          # logScale / logit range
    ]).astype(np.float32)
    got = oracle.expf(x)
    want = np.exp(x.astype(np.float64)).astype(np.float32)  # correctly rounded except ~2^-29 cases
    d = _ulp_diff(got, want)
    # subnormal outputs are judged in absolute terms
    normal = want >= np.finfo(np.float32).tiny
    assert d[normal].max() <= 1
    assert np.abs(got[~normal].astype(np.float64) - want[~normal].astype(np.float64)).max() <= 2 ** -149
    assert (d[normal] == 0).mean() > 0.9


def test_expf_special_values():
    x = np.array([0.0, -0.0, -200.0, -104.0, 89.0, 200.0, np.inf, -np.inf, np.nan], np.float32)
    y = oracle.expf(x)
    assert y[0] == 1.0 and y[1] == 1.0
    assert y[2] == 0.0 and y[3] == 0.0
    assert np.isinf(y[4]) and np.isinf(y[5]) and np.isinf(y[6])
    assert y[7] == 0.0
    assert np.isnan(y[8])


def test_expf_monotone_on_rasterize_range():
    x = np.arange(-5.0, 0.0, 1e-5, dtype=np.float64).astype(np.float32)
    y = oracle.expf(x)
    assert (np.diff(y) >= 0).all()


def test_float_to_sortable_preserves_order():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.standard_normal(100_000) * 10.0 ** rng.uniform(-30, 30, 100_000),
                        [0.0, -0.0, 1e-45, -1e-45, np.inf, -np.inf, 3.4e38, -3.4e38]]).astype(np.float32)
    keys = np.array([oracle.float_to_sortable(float(f)) for f in v], np.uint64)
    # -0.0 sorts strictly before +0.0 in the key space; both are equal floats
    order_k = np.argsort(keys, kind="stable")
    sv = v[order_k]
    assert (np.diff(sv.astype(np.float64)) >= 0).all()
    assert oracle.float_to_sortable(1.0) == 0xBF800000
    assert oracle.float_to_sortable(-1.0) == (~np.uint32(0xBF800000)) & 0xFFFFFFFF


def test_u32sat_matches_wgsl_conversion():
    assert oracle.u32sat(-0.5) == 0
    assert oracle.u32sat(-1e30) == 0
    assert oracle.u32sat(0.999) == 0
    assert oracle.u32sat(7.99) == 7
    assert oracle.u32sat(4294967296.0 * 2) == 0xFFFFFFFF
    assert oracle.u32sat(float("nan")) == 0


def test_focal_closed_form():
    for h, fov in [(720, math.pi / 3), (2160, math.pi / 3), (1080, 1.0), (100, 2.5)]:
        cam = S.default_camera(1280, h)
        cam.fov_y = fov
        assert oracle.focal(cam) == h / (2 * math.tan(fov / 2))


def test_morton_round_trip():
    seen = set()
    for m in range(64):
        x, y, z = oracle.morton_decode3(m)
        assert 0 <= x < 4 and 0 <= y < 4 and 0 <= z < 4
        assert oracle.morton_encode3(x, y, z) == m
        seen.add((x, y, z))
    assert len(seen) == 64
    # bit order z2 y2 x2 z1 y1 x1
    assert oracle.morton_encode3(1, 0, 0) == 1
    assert oracle.morton_encode3(0, 1, 0) == 2
    assert oracle.morton_encode3(0, 0, 1) == 4
    assert oracle.morton_encode3(2, 0, 0) == 8
    assert oracle.morton_encode3(0, 0, 2) == 32


def test_sort_pairs_is_stable_counting_sort():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 3000, 200_000).astype(np.uint32)
    vals = np.arange(keys.size, dtype=np.uint32)
    k, v = oracle.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("proj", ["pinhole", "equirect"])
def test_cull_and_depth_sort(proj):
    data = S.gaussian_cloud(50_000, seed=3, pos_std=2.0, pos_clip=6.0)
    cam = S.default_camera(640, 360) if proj == "pinhole" else S.equirect_camera(512, 256)
    if proj == "equirect":
        cam.near = 0.5
    idx, dep = oracle.cull(data, cam)
    p = data.positions.reshape(-1, 3).astype(np.float64)
    m = cam.view_matrix.astype(np.float64)
    cz = ((m[2] * p[:, 0] + m[6] * p[:, 1]) + m[10] * p[:, 2]) + m[14]
    if proj == "pinhole":
        depth = -cz
        vis = (depth > cam.near) & (depth < cam.far)
    else:
        cx = ((m[0] * p[:, 0] + m[4] * p[:, 1]) + m[8] * p[:, 2]) + m[12]
        cy = ((m[1] * p[:, 0] + m[5] * p[:, 1]) + m[9] * p[:, 2]) + m[13]
        r2 = (cx * cx + cy * cy) + cz * cz
        vis = r2 > cam.near * cam.near
        depth = np.sqrt(r2)
    assert 0 < vis.sum() < data.count  # both branches exercised
    np.testing.assert_array_equal(idx, np.nonzero(vis)[0])
    np.testing.assert_array_equal(dep, depth[vis].astype(np.float32))
    sorted_idx = oracle.depth_sort(idx, dep)
    # stable LSD radix on sortable keys == stable sort by f32 depth value
    np.testing.assert_array_equal(sorted_idx, idx[np.argsort(dep, kind="stable")])


def _round_f32(q):
    """Round a rational exactly to the nearest f32 (ties to even)."""
    from fractions import Fraction
    f = np.float32(float(q))
    best = None
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(cand):
            continue
        d = abs(Fraction(float(cand)) - q)
        key = (d, int(np.float32(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return best[1]


def test_fmaf_is_correctly_rounded():
    """The one fused operation of the rasterizer's specification: the oracle's
    fmaf equals a*b + c computed exactly and rounded once, including sums
    whose exponents are far apart (where an unfused or double-rounded
    evaluation differs)."""
    from fractions import Fraction
    rng = np.random.default_rng(17)
    a = rng.normal(0, 1, 3000).astype(np.float32)
    b = rng.normal(0, 1, 3000).astype(np.float32)
    c = (rng.normal(0, 1, 3000) * 10.0 ** rng.integers(-12, 12, 3000)).astype(np.float32)
    c[:500] = -(a[:500] * b[:500])            # near-cancellation
    got = oracle.fmaf(a, b, c)
    want = np.array([_round_f32(Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)))
                     for x, y, z in zip(a, b, c)], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    unfused = (a * b + c).astype(np.float32)
    assert (unfused.view(np.uint32) != want.view(np.uint32)).any()
