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
    """A long grid: more than 256 blocks along x (22 m at 0.08 m blocks),
    5 along y and 8 along z, so the pairs take the layer route with the
    layers along y."""
    data = _small_box_scene(6000, 5, box=((-11, 0, -0.3), (11, 0.4, 0.3)))
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert got.num_blocks_x > 256
    assert got.buffer.count > 1000


def test_voxel_grid_with_many_layers(device):
    """A 12 m column along z (150 blocks) a few blocks wide in x and y: the
    layer route takes y, the axis of fewest blocks, whatever the others
    hold."""
    data = _small_box_scene(4000, 6, box=((-0.2, 0, -6), (0.2, 0.3, 6)))
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert got.num_blocks_z > 140
    assert got.buffer.count > 500


@pytest.mark.parametrize("box", [((-0.15, 0, -1.5), (0.15, 3, 1.5)), ((-1.5, 0, -0.15), (1.5, 3, 0.15))],
                         ids=["layers_along_x", "layers_along_z"])
def test_voxel_layer_axis(device, box):
    """Slabs thin in x, then in z (10 blocks against 43-44): the layer route
    along x and along z, the key's other two axes ordered (y, z) and
    (x, y)."""
    data = _small_box_scene(5000, 8, box=box)
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    dims = (got.num_blocks_x, got.num_blocks_y, got.num_blocks_z)
    assert min(dims) <= 12 and max(dims) >= 40
    assert got.buffer.count > 500


def test_voxel_flat_route(device):
    """Gaussians along the diagonal of a 5.6 m cube: 76 blocks on every
    axis, more than the layer route's 64 layers, so the pairs are keyed by
    the flat block index and sorted on every bit."""
    rng = np.random.default_rng(7)
    n = 3000
    t = rng.random(n)[:, None]
    pos = (t * np.array([5.6, 5.6, 5.6]) + rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    ls = rng.uniform(math.log(0.005), math.log(0.08), (n, 3)).astype(np.float32)
    op = rng.normal(0, 2, n).astype(np.float32)
    data = SplatData(n, pos, q.astype(np.float32), ls, op, np.zeros(3 * n, np.float32))
    got, ref = assert_voxel_parity(device, data, 0.02, 0.1)
    assert min(got.num_blocks_x, got.num_blocks_y, got.num_blocks_z) > 64
    assert got.buffer.count > 1000
