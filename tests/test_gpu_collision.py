"""GPU parity of the voxel grid in HBM and the collision stages after the
voxelizer (fill, carve, query, mesh) against the CPU oracle
(oracle/collision.c). Bar: dense masks, block lists in insertion order, query
bits and mesh arrays (float32 positions and u32 indices) bit-exact."""
import numpy as np
import pytest

import oracle
from droplet_amd import _lib as L
from droplet_amd import synthetic as S
from droplet_amd.types import Bounds
from droplet_amd.voxel import (BlockMaskBuffer, VoxelGrid, VoxelizeOptions, VoxelizeResult,
                               voxelize, voxelize_resident)
from grids import from_volume, random_volume, serpentine, to_volume

pytestmark = pytest.mark.gpu


def upload(device, dense, nb, bmin=(0.0, 0.0, 0.0), res=1.0, shuffle_seed=None):
    idx, lo, hi = oracle.sparse_blocks(dense, nb)
    if shuffle_seed is not None:
        p = np.random.default_rng(shuffle_seed).permutation(idx.size)
        idx, lo, hi = idx[p], lo[p], hi[p]
    bmax = tuple(float(b + 4 * n * res) for b, n in zip(bmin, nb))
    r = VoxelizeResult(BlockMaskBuffer(idx, lo, hi), Bounds(tuple(map(float, bmin)), bmax), res,
                       float(nb[0]), float(nb[1]), float(nb[2]))
    return VoxelGrid.from_result(device, r)


def dense_of(grid: VoxelGrid, nb):
    r = grid.result()
    return oracle.dense_masks(r.buffer.block_idx, r.buffer.mask_lo, r.buffer.mask_hi, nb)


def assert_grid(grid: VoxelGrid, want_dense, nb, label):
    r = grid.result()
    idx, lo, hi = oracle.sparse_blocks(want_dense, nb)
    print(f"case: {label} nb={nb} blocks={idx.size} ms={grid.device_ms:.3f}", flush=True)
    assert np.array_equal(r.buffer.block_idx, idx)
    assert np.array_equal(r.buffer.mask_lo, lo)
    assert np.array_equal(r.buffer.mask_hi, hi)


CASES = [(0, (1, 1, 1), 0.3), (1, (9, 3, 5), 0.3), (2, (17, 9, 5), 0.45), (3, (8, 8, 4), 0.55),
         (4, (3, 20, 2), 0.25), (5, (12, 5, 13), 0.6)]


@pytest.mark.parametrize("seed,nb,p", CASES)
def test_upload_roundtrip_and_order(device, seed, nb, p):
    vol = random_volume(np.random.default_rng(seed), nb, p)
    dense, nb2 = from_volume(vol)
    with upload(device, dense, nb2, shuffle_seed=seed) as g:
        assert_grid(g, dense, nb2, "upload")
        v = g.device_view()
        assert v.num_blocks == nb2[0] * nb2[1] * nb2[2] and v.count == g.count
        assert tuple(v.nb) == nb2 and v.masks


@pytest.mark.parametrize("seed,nb,p", CASES)
def test_fill_matches_oracle(device, seed, nb, p):
    vol = random_volume(np.random.default_rng(seed), nb, p)
    dense, nb2 = from_volume(vol)
    with upload(device, dense, nb2) as g, g.fill() as f:
        assert_grid(f, oracle.vox_fill(dense, nb2), nb2, "fill")


@pytest.mark.parametrize("seed,nb,p", CASES)
def test_carve_matches_oracle(device, seed, nb, p):
    rng = np.random.default_rng(100 + seed)
    vol = random_volume(rng, nb, p)
    dense, nb2 = from_volume(vol)
    bmin, res = np.array([-2.5, 0.25, 7.0]), 0.05
    V = 4 * np.array(nb2)
    pts = bmin + rng.random((40, 3)) * V * res
    pts = np.concatenate([pts, [[np.nan, 0, 0], bmin - 1.0, bmin + V * res]])
    with upload(device, dense, nb2, bmin, res) as g, g.carve(pts) as c:
        assert_grid(c, oracle.vox_carve(dense, nb2, bmin, res, pts), nb2, "carve")


def test_carve_follows_a_long_corridor(device):
    """The corridor crosses tile borders many times, so the flood needs many
    launches before one changes nothing."""
    vol = serpentine((16, 1, 16))
    dense, nb = from_volume(vol)
    with upload(device, dense, nb) as g, g.carve([[1.5, 1.5, 1.5]]) as c:
        want = oracle.vox_carve(dense, nb, (0, 0, 0), 1.0, [[1.5, 1.5, 1.5]])
        assert_grid(c, want, nb, "corridor")
        assert int((~to_volume(dense_of(c, nb), nb)).sum()) == int((~vol).sum())


@pytest.mark.parametrize("seed,nb,p", CASES[:4])
def test_query_matches_oracle(device, seed, nb, p):
    rng = np.random.default_rng(200 + seed)
    vol = random_volume(rng, nb, p)
    dense, nb2 = from_volume(vol)
    bmin, res = np.array([0.5, -1.0, 3.25]), 0.125
    V = 4 * np.array(nb2)
    pts = bmin + (rng.random((5000, 3)) * 1.2 - 0.1) * V * res
    with upload(device, dense, nb2, bmin, res) as g:
        got = g.query(pts)
    want = oracle.vox_query(dense, nb2, bmin, res, pts).astype(bool)
    assert np.array_equal(got, want) and want.any() and not want.all()


@pytest.mark.parametrize("seed,nb,p", CASES)
def test_mesh_matches_oracle(device, seed, nb, p):
    vol = random_volume(np.random.default_rng(300 + seed), nb, p)
    dense, nb2 = from_volume(vol)
    bmin, res = (-1.75, 0.5, 2.0), 0.0371
    with upload(device, dense, nb2, bmin, res) as g:
        m = g.mesh()
    pos, idx = oracle.vox_mesh(dense, nb2, bmin, res)
    print(f"case: mesh nb={nb2} quads={m.quads} ms={m.device_ms:.3f}", flush=True)
    assert m.positions.shape == pos.shape and np.array_equal(m.positions, pos)
    assert np.array_equal(m.indices, idx)


def test_mesh_rows_longer_than_a_word(device):
    """Runs that cross 64-bit word borders (V along u > 64)."""
    vol = np.zeros((4 * 40, 8, 12), bool)
    vol[3:150, 2:6, 1:11] = True
    vol[64, 3, 5] = False
    vol[127:129, 4, 2] = False
    dense, nb = from_volume(vol)
    with upload(device, dense, nb) as g:
        m = g.mesh()
    pos, idx = oracle.vox_mesh(dense, nb, (0, 0, 0), 1.0)
    assert np.array_equal(m.positions, pos) and np.array_equal(m.indices, idx)


def test_voxelize_fill_carve_mesh_pipeline(device):
    data, cfg = S.voxel_workload(20000, seed=3)
    opts = VoxelizeOptions(0.05, cfg["opacity_cutoff"])
    with voxelize_resident(device, data, opts) as g:
        r = g.result()
        nb = (int(r.num_blocks_x), int(r.num_blocks_y), int(r.num_blocks_z))
        dense = oracle.dense_masks(r.buffer.block_idx, r.buffer.mask_lo, r.buffer.mask_hi, nb)
        bmin = r.bounds.min
        with g.fill() as f:
            want_fill = oracle.vox_fill(dense, nb)
            assert_grid(f, want_fill, nb, "pipeline fill")
            m = f.mesh()
            pos, idx = oracle.vox_mesh(want_fill, nb, bmin, 0.05)
            print(f"case: pipeline mesh quads={m.quads} ms={m.device_ms:.3f}", flush=True)
            assert np.array_equal(m.positions, pos) and np.array_equal(m.indices, idx)
        seed_pt = [[float(bmin[0]) + 0.01, float(bmin[1]) + 0.01, float(bmin[2]) + 0.01]]
        with g.carve(seed_pt) as c:
            assert_grid(c, oracle.vox_carve(dense, nb, bmin, 0.05, seed_pt), nb, "pipeline carve")
    # the host result of voxelize equals the resident grid's
    r2 = voxelize(device, data, opts)
    assert np.array_equal(r2.buffer.block_idx, r.buffer.block_idx)
    assert np.array_equal(r2.buffer.mask_lo, r.buffer.mask_lo)


def test_repeated_blocks_or_merge(device):
    r = VoxelizeResult(BlockMaskBuffer(np.array([3, 1, 3, 3], np.uint32), np.array([1, 0, 4, 0], np.uint32),
                                       np.array([0, 0, 0, 2], np.uint32)),
                       Bounds((0.0, 0.0, 0.0), (8.0, 4.0, 4.0)), 1.0, 2.0, 1.0, 2.0)
    with VoxelGrid.from_result(device, r) as g:
        out = g.result()
    # block 1 has an empty mask and is dropped, as addBlock drops it
    assert out.buffer.block_idx.tolist() == [3]
    assert out.buffer.mask_lo.tolist() == [5] and out.buffer.mask_hi.tolist() == [2]


def test_empty_grids(device):
    with voxelize_resident(device, S.empty_scene(), VoxelizeOptions(0.05, 0.1)) as g:
        assert g.count == 0
        assert g.mesh().quads == 0
        with g.fill() as f:
            assert f.count == 0
        assert not g.query([[0.0, 0.0, 0.0]]).any()
    dense = np.zeros(2 * 3 * 1, np.uint64)
    with upload(device, dense, (2, 3, 1)) as g:
        assert g.count == 0 and g.mesh().quads == 0
        with g.fill() as f:
            assert f.count == 0
        with g.carve(np.zeros((0, 3))) as c:
            assert c.count == 6 and c.result().buffer.popcount() == 6 * 64


def test_invalid_uploads_are_rejected(device):
    base = dict(bounds=Bounds((0.0, 0.0, 0.0), (4.0, 4.0, 4.0)), voxel_resolution=1.0, num_blocks_x=1.0,
                num_blocks_y=1.0, num_blocks_z=1.0)
    bad_idx = VoxelizeResult(BlockMaskBuffer(np.array([1], np.uint32), np.array([1], np.uint32),
                                             np.array([0], np.uint32)), **base)
    with pytest.raises(L.DropletError):
        VoxelGrid.from_result(device, bad_idx)
    frac = VoxelizeResult(BlockMaskBuffer(), **{**base, "num_blocks_x": 1.5})
    with pytest.raises(L.DropletError):
        VoxelGrid.from_result(device, frac)
    big = VoxelizeResult(BlockMaskBuffer(), **{**base, "num_blocks_x": 65536.0, "num_blocks_y": 65536.0})
    with pytest.raises(L.DropletError):
        VoxelGrid.from_result(device, big)
