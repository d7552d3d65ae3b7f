"""Pins the oracle's voxelizer on the CPU.

* a single isotropic Gaussian: the solid voxels are exactly those where
  1 - exp(-o e^{-d^2/2s^2}) >= cutoff, d = distance from the centre to the
  closest point of the voxel cube (f64 closed form, apart from voxels whose
  alpha sits within 1e-5 of the cutoff);
* AABB extents against R diag(s) in f64 (aabb.ts:22-62);
* BVH candidate lists are complete: the sum of batch list lengths equals the
  brute-force count of (batch box, Gaussian AABB) overlaps;
* the BVH order is a permutation of all Gaussians;
* empty scene: 0 blocks, numBlocks = -Infinity (voxelize.ts:74-76 on
  +/-Infinity bounds);
* the browser test page voxel cases (examples/browser/test.ts:253-274).
"""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.types import SplatData


def _sig(x):
    return 1 / (1 + math.exp(-x))


def _decode(res, nbx, nby):
    out = {}
    for b, lo, hi in zip(res["block_idx"], res["mask_lo"], res["mask_hi"]):
        out[int(b)] = (int(lo) | (int(hi) << 32))
    return out


@pytest.mark.parametrize("ls,logit,cut", [(-2.5, 2.0, 0.1), (-3.0, 0.0, 0.3), (-2.0, -1.0, 0.05)])
def test_single_gaussian_ball_closed_form(ls, logit, cut):
    p = np.array([0.113, -0.071, 0.052], np.float32)
    data = SplatData(1, p, [1, 0, 0, 0], [ls] * 3, [logit], [0, 0, 0])
    res_v = 0.02
    r = oracle.voxelize(data, voxel_resolution=res_v, opacity_cutoff=cut, near_eps=1e-5)
    nb = [int(v) for v in r["num_blocks_xyz"]]
    bmn = np.array(r["bounds_min"])
    got = _decode(r, nb[0], nb[1])
    s = math.exp(float(np.float32(ls)))
    ext = float(np.float32(3 * s))
    o = _sig(float(np.float32(logit)))
    half = res_v / 2
    near = set(zip(r["near_block"].tolist(), r["near_bit"].tolist()))
    mism = 0
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                flat = bx + by * nb[0] + bz * nb[0] * nb[1]
                mask = got.get(flat, 0)
                for lin in range(64):
                    lx, ly, lz = lin & 3, (lin >> 2) & 3, lin >> 4
                    vc = bmn + (np.array([bx, by, bz]) * 4 + np.array([lx, ly, lz]) + 0.5) * res_v
                    inside = np.all(np.abs(vc - p) <= ext + half)
                    closest = np.clip(p, vc - half, vc + half)
                    d2 = float(np.sum((closest - p) ** 2)) / (s * s)
                    alpha = 1 - math.exp(-o * math.exp(-0.5 * d2)) if inside else 0.0
                    want = alpha >= cut
                    have = bool((mask >> lin) & 1)
                    if want != have:
                        assert (flat, lin) in near or abs(alpha - cut) < 1e-5, (flat, lin, alpha)
                        mism += 1
    assert mism <= 2
    assert r["block_idx"].size > 0


def _rot(q):
    """Rotation matrix of a unit quaternion [w, x, y, z] in f64."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("q,ls", [
    ((0.9, 0.2, -0.3, 0.25), (math.log(0.12), math.log(0.03), math.log(0.06))),
    ((0.5, -0.6, 0.4, 0.45), (math.log(0.02), math.log(0.1), math.log(0.04))),
    ((math.cos(0.3), 0.0, 0.0, math.sin(0.3)), (math.log(0.15), math.log(0.025), math.log(0.025))),
])
def test_rotated_anisotropic_gaussian_closed_form(q, ls):
    """A single rotated, anisotropic Gaussian (voxel/shaders/voxelize.ts:62-84):
    the solid voxels are those where 1 - exp(-o exp(-d2/2)) >= cutoff with
    d2 = |diag(1/s) R(q)^T (closest - g)|^2, evaluated in f64 from the
    rotation matrix of q, apart from voxels within 1e-5 of the cutoff. The
    inverse rotation R^T is what the shader's conjugated quaternion applies;
    the test also checks that R instead of R^T would change many voxels, so
    a dropped conjugate cannot pass."""
    qn = np.array(q, np.float64)
    qn = (qn / np.linalg.norm(qn)).astype(np.float32)
    p = np.array([0.031, -0.017, 0.044], np.float32)
    logit, cut, res_v = 2.0, 0.1, 0.02
    data = SplatData(1, p, qn, np.array(ls, np.float32), [logit], [0, 0, 0])
    r = oracle.voxelize(data, voxel_resolution=res_v, opacity_cutoff=cut, near_eps=1e-5)
    nb = [int(v) for v in r["num_blocks_xyz"]]
    bmn = np.array(r["bounds_min"])
    got = _decode(r, nb[0], nb[1])
    R = _rot(qn.astype(np.float64))
    s = np.exp(np.array(ls, np.float32).astype(np.float64))
    ext = oracle.aabb_extents(data).astype(np.float64)
    o = _sig(float(np.float32(logit)))
    half = res_v / 2
    pd = p.astype(np.float64)
    # every voxel centre of the grid
    bz, by, bx, lz, ly, lx = np.meshgrid(np.arange(nb[2]), np.arange(nb[1]), np.arange(nb[0]),
                                         np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    flat = (bx + by * nb[0] + bz * nb[0] * nb[1]).ravel()
    lin = (lx + 4 * ly + 16 * lz).ravel()
    vc = bmn + (np.stack([bx * 4 + lx, by * 4 + ly, bz * 4 + lz], -1).reshape(-1, 3) + 0.5) * res_v
    inside = np.all(np.abs(vc - pd) <= ext + half, axis=1)
    cd = np.clip(pd, vc - half, vc + half) - pd

    def alpha_of(M):
        local = cd @ M.T              # rows: M @ cd
        d2 = np.sum((local / s) ** 2, axis=1)
        return np.where(inside, 1 - np.exp(-o * np.exp(-0.5 * d2)), 0.0)

    want_a = alpha_of(R.T)
    wrong_a = alpha_of(R)
    masks = np.array([got.get(int(f), 0) for f in flat], dtype=object)
    have = np.array([(int(m) >> int(b)) & 1 for m, b in zip(masks, lin)], bool)
    want = want_a >= cut
    near = set(zip(r["near_block"].tolist(), r["near_bit"].tolist()))
    bad = np.nonzero(want != have)[0]
    for k in bad:
        assert (int(flat[k]), int(lin[k])) in near or abs(want_a[k] - cut) < 1e-5, (flat[k], lin[k], want_a[k])
    assert bad.size <= 2
    assert want.sum() > 100
    # the case separates R^T from R
    assert ((wrong_a >= cut) != want).sum() > 20


def test_aabb_extents_against_rotation_matrix():
    data = S.gaussian_cloud(2000, seed=4)
    ext = oracle.aabb_extents(data).reshape(-1, 3).astype(np.float64)
    q = data.rotations.reshape(-1, 4).astype(np.float64)
    s = np.exp(data.log_scales.reshape(-1, 3).astype(np.float64))
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)
    want = 3 * np.sqrt(np.einsum("nij,nj->ni", R ** 2, s ** 2))
    np.testing.assert_allclose(ext, want, rtol=1e-6)
    # identity rotation: ext = 3 s exactly
    one = SplatData(1, [0, 0, 0], [1, 0, 0, 0], [-2.0, -3.0, -1.5], [0], [0, 0, 0])
    np.testing.assert_array_equal(oracle.aabb_extents(one), np.float32(3 * np.exp([-2.0, -3.0, -1.5])))


def test_bvh_order_is_permutation_and_lists_are_complete():
    data = S.gaussian_cloud(3000, seed=5, pos_std=0.4, log_scale_range=(math.log(0.005), math.log(0.04)))
    ext = oracle.aabb_extents(data)
    perm = oracle.bvh_order(data.positions, ext)
    assert sorted(perm.tolist()) == list(range(data.count))
    r = oracle.voxelize(data, voxel_resolution=0.02, opacity_cutoff=0.1)
    bmn = np.array(r["bounds_min"])
    nb = [int(v) for v in r["num_blocks_xyz"]]
    bs = 4 * 0.02
    p = data.positions.reshape(-1, 3).astype(np.float64)
    e = ext.reshape(-1, 3).astype(np.float64)
    total, batches = 0, 0
    for bz in range(0, nb[2], 16):
        for by in range(0, nb[1], 16):
            for bx in range(0, nb[0], 16):
                c = [min(16, nb[0] - bx), min(16, nb[1] - by), min(16, nb[2] - bz)]
                qmn = np.array([bmn[0] + bx * bs, bmn[1] + by * bs, bmn[2] + bz * bs])
                qmx = qmn + np.array(c) * bs
                ok = ~(((p + e) < qmn) | ((p - e) > qmx)).any(1)
                total += int(ok.sum())
                batches += int(ok.any())
    assert r["num_candidates"] == total
    assert r["num_batches"] == batches


def test_empty_scene():
    r = oracle.voxelize(S.empty_scene(), voxel_resolution=0.05, opacity_cutoff=0.1)
    assert r["block_idx"].size == 0
    assert all(v == -math.inf for v in r["num_blocks_xyz"])


def test_browser_sphere_voxelizes():
    r = oracle.voxelize(S.many_splats(5000, "sphere"), voxel_resolution=0.05, opacity_cutoff=0.1)
    assert r["block_idx"].size > 0
    # insertion order: block ids inside one batch ascend in (lz, ly, lx) order
    assert len(set(r["block_idx"].tolist())) == r["block_idx"].size


def test_reference_test_count_known_answers():
    """ref_tests counts the (voxel, candidate) AABB tests the reference runs
    (each voxel walks its batch list until sigma >= 7). One Gaussian never
    reaches 7, so every voxel of its one batch tests it once; with opaque
    copies stacked on one point the centre voxels stop early, so the count
    falls below voxels x candidates but stays above the passing tests."""
    from droplet_amd.types import Bounds
    one = SplatData(1, [0.0, 0.0, 0.0], [1, 0, 0, 0], [-3.0] * 3, [2.0], [0, 0, 0])
    # a grid wider than the Gaussian's AABB: the outer voxels fail the test
    r = oracle.voxelize(one, voxel_resolution=0.02, opacity_cutoff=0.1, bounds=Bounds((-0.32,) * 3, (0.32,) * 3))
    nb = [int(v) for v in r["num_blocks_xyz"]]
    assert nb == [8, 8, 8] and r["num_batches"] == 1
    assert r["ref_tests"] == 64 * 8 ** 3
    # the passing tests: voxels whose centre lies within ext + half of the centre per axis
    ext, half = float(np.float32(3 * math.exp(-3.0))), 0.01
    c = -0.32 + (np.arange(32) + 0.5) * 0.02
    inside = int((np.abs(c) <= ext + half).sum())
    assert r["contrib_pairs"] == inside ** 3 < r["ref_tests"]
    k = 12
    stack = SplatData(k, np.zeros(3 * k, np.float32), np.tile([1, 0, 0, 0], k).astype(np.float32),
                      np.full(3 * k, -3.0, np.float32), np.full(k, 9.0, np.float32), np.zeros(3 * k, np.float32))
    r = oracle.voxelize(stack, voxel_resolution=0.02, opacity_cutoff=0.1)
    nb = [int(v) for v in r["num_blocks_xyz"]]
    full = 64 * nb[0] * nb[1] * nb[2] * k
    assert r["num_batches"] == 1 and r["num_candidates"] == k
    assert r["contrib_pairs"] <= r["ref_tests"] < full
