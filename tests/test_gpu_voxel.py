"""GPU parity of the voxelizer: libdroplet_hip.so against the CPU oracle.

Bar (from the task): block lists (indices, insertion order) and 64-bit masks
bit-exact, except voxels whose alpha lies within 1e-6 of the cutoff; those are
counted, printed and must be listed by the oracle as near-cutoff. The device
sums a voxel's Gaussians in index order, the reference in BVH order, so only
such near-cutoff voxels can differ.
"""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.types import Bounds, SplatData
from droplet_amd.voxel import VoxelizeOptions, voxelize

pytestmark = pytest.mark.gpu

NEAR_EPS = 1e-6


def _small_box_scene(n, seed, box=((-1, 0, -1), (1, 0.6, 1))):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(box[0], float), np.array(box[1], float)
    pos = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    ls = rng.uniform(math.log(0.005), math.log(0.08), (n, 3))
    op = rng.normal(0, 2, n)
    return SplatData(n, pos, q.astype(np.float32), ls.astype(np.float32), op.astype(np.float32),
                     np.zeros(3 * n, np.float32))


def assert_voxel_parity(device, data, res, cut, bounds=None):
    got = voxelize(device, data, VoxelizeOptions(res, cut, bounds))
    ref = oracle.voxelize(data, voxel_resolution=res, opacity_cutoff=cut, bounds=bounds, near_eps=NEAR_EPS)
    print(f"case: n={data.count} res={res} cut={cut} blocks={ref['block_idx'].size} near={ref['num_near']} "
          f"gpu_ms={got.device_ms:.2f}", flush=True)
    assert tuple(got.bounds.min) == ref["bounds_min"]
    assert tuple(got.bounds.max) == ref["bounds_max"]
    assert (got.num_blocks_x, got.num_blocks_y, got.num_blocks_z) == ref["num_blocks_xyz"]
    near = {}
    for b, bit in zip(ref["near_block"].tolist(), ref["near_bit"].tolist()):
        near.setdefault(b, 0)
        near[b] |= 1 << bit
    g = {int(b): (int(lo) | int(hi) << 32) for b, lo, hi in
         zip(got.buffer.block_idx, got.buffer.mask_lo, got.buffer.mask_hi)}
    r = {int(b): (int(lo) | int(hi) << 32) for b, lo, hi in zip(ref["block_idx"], ref["mask_lo"], ref["mask_hi"])}
    diff_bits = 0
    for b in set(g) | set(r):
        d = g.get(b, 0) ^ r.get(b, 0)
        if d:
            assert d & ~near.get(b, 0) == 0, f"block {b}: differing bits outside the near-cutoff set"
            diff_bits += bin(d).count("1")
    print(f"near-cutoff voxels that differ: {diff_bits} (of {ref['num_near']} near-cutoff)", flush=True)
    # insertion order: identical wherever both hold the block
    common = [b for b in ref["block_idx"].tolist() if b in g]
    order_g = [b for b in got.buffer.block_idx.tolist() if b in r]
    assert common == order_g
    if diff_bits == 0:
        np.testing.assert_array_equal(got.buffer.block_idx, ref["block_idx"])
        np.testing.assert_array_equal(got.buffer.mask_lo, ref["mask_lo"])
        np.testing.assert_array_equal(got.buffer.mask_hi, ref["mask_hi"])
    return got, ref


def test_voxel_small_scene_parity(device):
    assert_voxel_parity(device, _small_box_scene(20_000, 0), 0.02, 0.1)


def test_voxel_sparse_scene_coarse(device):
    assert_voxel_parity(device, _small_box_scene(4000, 1, box=((-3, -1, -2), (3, 1, 2))), 0.05, 0.3)


def test_voxel_sphere_and_bounds_override(device):
    data = S.many_splats(5000, "sphere")
    got, ref = assert_voxel_parity(device, data, 0.05, 0.1)
    assert got.buffer.count > 0  # test.ts:253-260
    b = Bounds((-0.6, -0.4, -0.8), (0.9, 1.0, 0.5))
    assert_voxel_parity(device, data, 0.05, 0.1, bounds=b)


def test_voxel_empty_scene(device):
    got = voxelize(device, S.empty_scene(), VoxelizeOptions(0.05, 0.1))
    assert got.buffer.count == 0  # test.ts:262-274
    assert got.num_blocks_x == -math.inf


def test_voxel_single_gaussian(device):
    data = SplatData(1, [0.113, -0.071, 0.052], [1, 0, 0, 0], [-2.5] * 3, [2.0], [0, 0, 0])
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert got.buffer.popcount() > 100


def test_voxel_grid_wider_than_256_blocks(device):
    """A long grid: more than 256 blocks along x (22 m at 0.08 m blocks)
    and a few block layers, so the pairs take the layer route."""
    data = _small_box_scene(6000, 5, box=((-11, 0, -0.3), (11, 0.4, 0.3)))
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert got.num_blocks_x > 256
    assert got.buffer.count > 1000


def test_voxel_grid_with_many_layers(device):
    """Many block layers along z (a 12 m column at 0.08 m blocks: 150
    layers of a few blocks each): more than 64 layers, so the pairs take the
    flat route (keyed by the flat block index, sorted on every bit)."""
    data = _small_box_scene(4000, 6, box=((-0.2, 0, -6), (0.2, 0.3, 6)))
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert got.num_blocks_z > 140
    assert got.buffer.count > 500
