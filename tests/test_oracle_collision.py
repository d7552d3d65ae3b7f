"""Pins the CPU oracle of the collision stages (oracle/collision.c): fill,
carve, query and mesh over a voxel grid. The reference has no code for these
stages (README.md:108 names them as downstream grid arithmetic on
BlockMaskBuffer), so the definitions are this library's; they are pinned here
by scipy.ndimage (hole filling and connected-component labelling with
6-connectivity), by closed forms (face sets from shifted XORs, quad coverage)
and by known answers on tiny grids."""
import numpy as np
import pytest
from scipy import ndimage

import oracle
from grids import from_volume, random_volume, serpentine, to_volume

CONN6 = ndimage.generate_binary_structure(3, 1)


def test_volume_roundtrip():
    rng = np.random.default_rng(1)
    vol = rng.random((8, 12, 4)) < 0.5
    m, nb = from_volume(vol)
    assert nb == (2, 3, 1)
    assert np.array_equal(to_volume(m, nb), vol)
    # bit layout: voxel (1, 2, 3) of block 0 is bit 1 + 8 + 48
    m = np.zeros(1, np.uint64)
    m[0] = np.uint64(1) << np.uint64(1 + 4 * 2 + 16 * 3)
    assert np.argwhere(to_volume(m, (1, 1, 1))).tolist() == [[1, 2, 3]]


@pytest.mark.parametrize("seed,nb,p", [(0, (1, 1, 1), 0.3), (1, (3, 2, 4), 0.25), (2, (5, 4, 3), 0.45),
                                       (3, (6, 3, 5), 0.6), (4, (2, 7, 2), 0.35)])
def test_fill_matches_binary_fill_holes(seed, nb, p):
    rng = np.random.default_rng(seed)
    vol = random_volume(rng, nb, p)
    m, nb2 = from_volume(vol)
    got = to_volume(oracle.vox_fill(m, nb2), nb2)
    want = ndimage.binary_fill_holes(vol, structure=CONN6)
    assert np.array_equal(got, want)
    assert (got & ~vol).any() or not (want & ~vol).any()


def test_fill_closes_a_hollow_shell():
    vol = np.zeros((12, 12, 12), bool)
    vol[2:10, 2:10, 2:10] = True
    vol[3:9, 3:9, 3:9] = False
    m, nb = from_volume(vol)
    got = to_volume(oracle.vox_fill(m, nb), nb)
    assert got.sum() == 8 ** 3
    # a single hole in the shell lets the outside in: nothing to fill
    vol[5, 5, 2] = False
    m, nb = from_volume(vol)
    assert np.array_equal(to_volume(oracle.vox_fill(m, nb), nb), vol)


@pytest.mark.parametrize("seed,nb,p", [(5, (3, 3, 3), 0.35), (6, (4, 2, 5), 0.5), (7, (1, 1, 1), 0.2)])
def test_carve_matches_component_labels(seed, nb, p):
    rng = np.random.default_rng(seed)
    vol = random_volume(rng, nb, p)
    m, nb2 = from_volume(vol)
    bmin = np.array([-1.25, 0.5, 2.0])
    res = 0.125
    V = np.array(vol.shape)
    vox = rng.integers(0, V, size=(12, 3))
    pts = bmin + (vox + rng.random((12, 3))) * res
    # points outside the grid and NaN seed nothing
    pts = np.concatenate([pts, [[bmin[0] - 1e-9, bmin[1], bmin[2]], [np.nan, 0, 0],
                                bmin + V * res, bmin + V * res - 1e-9]])
    got = to_volume(oracle.vox_carve(m, nb2, bmin, res, pts), nb2)
    labels, _ = ndimage.label(~vol, structure=CONN6)
    seeds = []
    for q in pts:
        f = np.floor((q - bmin) / res)
        if np.all(f >= 0) and np.all(f < V):
            seeds.append(labels[int(f[0]), int(f[1]), int(f[2])])
    seeds = [s for s in seeds if s > 0]
    reach = np.isin(labels, seeds) & (labels > 0)
    assert np.array_equal(got, ~reach)


def test_carve_without_seeds_is_all_solid():
    vol = np.zeros((4, 8, 4), bool)
    m, nb = from_volume(vol)
    got = oracle.vox_carve(m, nb, (0, 0, 0), 1.0, np.zeros((0, 3)))
    assert np.all(got == np.uint64(0xFFFFFFFFFFFFFFFF))


def test_query_known_answers():
    rng = np.random.default_rng(9)
    vol = rng.random((8, 4, 12)) < 0.5
    m, nb = from_volume(vol)
    bmin, res = np.array([0.5, -2.0, 1.0]), 0.25
    V = np.array(vol.shape)
    vox = rng.integers(0, V, size=(200, 3))
    pts = bmin + (vox + 0.5) * res
    got = oracle.vox_query(m, nb, bmin, res, pts)
    assert np.array_equal(got.astype(bool), vol[vox[:, 0], vox[:, 1], vox[:, 2]])
    out = oracle.vox_query(m, nb, bmin, res, np.array([bmin - 0.01, bmin + V * res, [np.nan] * 3]))
    assert out.tolist() == [0, 0, 0]


# --------------------------------------------------------------------------
# mesh
# --------------------------------------------------------------------------
U_AXIS, V_AXIS = (1, 0, 0), (2, 2, 1)


def face_sets(vol):
    """F[d][x, y, z]: solid voxel whose neighbour along d is empty/outside."""
    pad = np.pad(vol, 1)
    out = []
    for d in range(6):
        a, sg = d // 2, (1 if d % 2 == 0 else -1)
        nbr = np.roll(pad, -sg, axis=a)[1:-1, 1:-1, 1:-1]
        out.append(vol & ~nbr)
    return out


def decode_quads(pos, idx):
    """(d, s, u0, u1, v0, v1) of every quad from integer-valued vertices."""
    q = pos.reshape(-1, 4, 3).astype(np.float64)
    tri = idx.reshape(-1, 2, 3)
    quads = []
    for k in range(q.shape[0]):
        c = q[k]
        a = int(np.where(np.all(c == c[0], axis=0))[0][0])
        ua, va = U_AXIS[a], V_AXIS[a]
        p0, p1, p2 = pos[tri[k, 0]].astype(np.float64)
        n = np.cross(p1 - p0, p2 - p0)
        assert n[a] != 0 and n[ua] == 0 and n[va] == 0
        positive = n[a] > 0
        # the second triangle faces the same way
        r0, r1, r2 = pos[tri[k, 1]].astype(np.float64)
        assert np.sign(np.cross(r1 - r0, r2 - r0)[a]) == np.sign(n[a])
        d = 2 * a + (0 if positive else 1)
        plane = int(c[0, a])
        s = plane - 1 if positive else plane
        u0, u1 = int(c[:, ua].min()), int(c[:, ua].max())
        v0, v1 = int(c[:, va].min()), int(c[:, va].max())
        # vertex order (u0,v0) (u1,v0) (u1,v1) (u0,v1)
        assert [(int(x[ua]), int(x[va])) for x in c] == [(u0, v0), (u1, v0), (u1, v1), (u0, v1)]
        assert set(tri[k].ravel().tolist()) == {4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 3}
        quads.append((d, s, u0, u1, v0, v1))
    return quads


def plane_view(F, d):
    """F_d as [s, u, v]."""
    a = d // 2
    return np.moveaxis(F, (a, U_AXIS[a], V_AXIS[a]), (0, 1, 2))


def exact_run(row, u0, u1):
    return (row[u0:u1].all() and (u0 == 0 or not row[u0 - 1]) and (u1 == row.size or not row[u1]))


def check_mesh(vol, pos, idx):
    F = face_sets(vol)
    quads = decode_quads(pos, idx)
    cover = [np.zeros(plane_view(F[d], d).shape, np.int32) for d in range(6)]
    keys = []
    for d, s, u0, u1, v0, v1 in quads:
        P = plane_view(F[d], d)
        cover[d][s, u0:u1, v0:v1] += 1
        for v in range(v0, v1):
            assert exact_run(P[s, :, v], u0, u1)
        assert v0 == 0 or not exact_run(P[s, :, v0 - 1], u0, u1)
        assert v1 == P.shape[2] or not exact_run(P[s, :, v1], u0, u1)
        keys.append((d, s, v0, u0))
    for d in range(6):
        assert np.array_equal(cover[d], plane_view(F[d], d).astype(np.int32))
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    return quads


def test_mesh_single_voxel_known_answer():
    m = np.array([1], np.uint64)
    pos, idx = oracle.vox_mesh(m, (1, 1, 1), (0, 0, 0), 1.0)
    assert pos.shape == (24, 3) and idx.shape == (12, 3)
    assert pos[:4].tolist() == [[1, 0, 0], [1, 1, 0], [1, 1, 1], [1, 0, 1]]
    assert idx[:4].tolist() == [[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]]


def test_mesh_full_block_is_six_quads():
    m = np.array([0xFFFFFFFFFFFFFFFF], np.uint64)
    pos, idx = oracle.vox_mesh(m, (1, 1, 1), (0, 0, 0), 1.0)
    quads = check_mesh(to_volume(m, (1, 1, 1)), pos, idx)
    assert [(q[2], q[3], q[4], q[5]) for q in quads] == [(0, 4, 0, 4)] * 6


@pytest.mark.parametrize("seed,nb,p", [(11, (2, 2, 2), 0.3), (12, (3, 1, 2), 0.6), (13, (2, 3, 3), 0.8),
                                       (14, (4, 2, 1), 0.05)])
def test_mesh_invariants(seed, nb, p):
    rng = np.random.default_rng(seed)
    vol = random_volume(rng, nb, p)
    m, nb2 = from_volume(vol)
    pos, idx = oracle.vox_mesh(m, nb2, (0, 0, 0), 1.0)
    quads = check_mesh(vol, pos, idx)
    area = sum((u1 - u0) * (v1 - v0) for _, _, u0, u1, v0, v1 in quads)
    assert area == sum(int(f.sum()) for f in face_sets(vol))


def test_mesh_positions_follow_bounds_and_resolution():
    rng = np.random.default_rng(15)
    vol = random_volume(rng, (2, 2, 3), 0.4)
    m, nb = from_volume(vol)
    pos_i, idx_i = oracle.vox_mesh(m, nb, (0, 0, 0), 1.0)
    bmin, res = np.array([-3.7, 0.123, 1e3]), 0.0371
    pos, idx = oracle.vox_mesh(m, nb, bmin, res)
    assert np.array_equal(idx, idx_i)
    want = (bmin + pos_i.astype(np.float64) * res).astype(np.float32)
    assert np.array_equal(pos, want)


def test_mesh_empty_grid():
    pos, idx = oracle.vox_mesh(np.zeros(6, np.uint64), (1, 2, 3), (0, 0, 0), 1.0)
    assert pos.shape == (0, 3) and idx.shape == (0, 3)


def test_mesh_after_fill_has_no_inner_faces():
    vol = np.zeros((8, 8, 8), bool)
    vol[1:7, 1:7, 1:7] = True
    vol[2:6, 2:6, 2:6] = False
    m, nb = from_volume(vol)
    pos, idx = oracle.vox_mesh(oracle.vox_fill(m, nb), nb, (0, 0, 0), 1.0)
    quads = check_mesh(to_volume(oracle.vox_fill(m, nb), nb), pos, idx)
    assert len(quads) == 6   # a solid 6^3 cube


def test_carve_follows_a_long_corridor():
    vol = serpentine((6, 1, 6))
    m, nb = from_volume(vol)
    free = int((~vol).sum())
    got = to_volume(oracle.vox_carve(m, nb, (0, 0, 0), 1.0, [[1.5, 1.5, 1.5]]), nb)
    assert int((~got).sum()) == free          # the whole corridor is one component
    labels, n = ndimage.label(~vol, structure=CONN6)
    assert n == 1


def test_insertion_order_matches_the_voxelizer():
    """block_insertion_order restates voxelize.ts's batch loops; the oracle
    voxelizer adds blocks through BlockMaskBuffer in that order."""
    from droplet_amd import synthetic as S
    data = S.gaussian_cloud(3000, seed=4, pos_std=0.6)
    ref = oracle.voxelize(data, voxel_resolution=0.04, opacity_cutoff=0.1)
    nb = tuple(int(v) for v in ref["num_blocks_xyz"])
    assert max(nb) > 16                       # more than one batch per axis
    order = oracle.block_insertion_order(nb)
    assert sorted(order.tolist()) == list(range(nb[0] * nb[1] * nb[2]))
    rank = np.empty(order.size, np.int64)
    rank[order] = np.arange(order.size)
    r = rank[ref["block_idx"].astype(np.int64)]
    assert np.all(np.diff(r) > 0)
    dense = oracle.dense_masks(ref["block_idx"], ref["mask_lo"], ref["mask_hi"], nb)
    idx, lo, hi = oracle.sparse_blocks(dense, nb)
    assert np.array_equal(idx, ref["block_idx"]) and np.array_equal(lo, ref["mask_lo"])
    assert np.array_equal(hi, ref["mask_hi"])
