"""Voxel grids for the collision-stage tests: conversion between a bool
volume [x, y, z] and the dense block grid (one 64-bit mask per 4x4x4 block at
the flat index bx + by*nb0 + bz*nb0*nb1, voxel bit x + 4y + 16z), and
generators of volumes with cavities and long corridors."""
import numpy as np


def to_volume(mask, nb):
    """bool[x, y, z] of a dense block grid (bit = x + 4y + 16z)."""
    nb0, nb1, nb2 = nb
    bits = ((np.asarray(mask, np.uint64)[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)
    v = bits.reshape(nb2, nb1, nb0, 4, 4, 4)            # bz by bx lz ly lx
    v = v.transpose(0, 3, 1, 4, 2, 5).reshape(4 * nb2, 4 * nb1, 4 * nb0)
    return np.ascontiguousarray(v.transpose(2, 1, 0))


def from_volume(vol):
    X, Y, Z = vol.shape
    nb = (X // 4, Y // 4, Z // 4)
    v = vol.transpose(2, 1, 0).reshape(nb[2], 4, nb[1], 4, nb[0], 4).transpose(0, 2, 4, 1, 3, 5)
    bits = v.reshape(-1, 64).astype(np.uint64)
    return (bits << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64), nb


def random_volume(rng, nb, p):
    vol = rng.random((4 * nb[0], 4 * nb[1], 4 * nb[2])) < p
    # plant hollow boxes so that fill has cavities to close
    for _ in range(3):
        lo = [int(rng.integers(0, max(1, 4 * n - 5))) for n in nb]
        hi = [min(4 * n, l + int(rng.integers(3, 7))) for l, n in zip(lo, nb)]
        if any(h - l < 3 for l, h in zip(lo, hi)):
            continue
        box = np.zeros_like(vol)
        box[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
        inner = np.zeros_like(vol)
        inner[lo[0] + 1:hi[0] - 1, lo[1] + 1:hi[1] - 1, lo[2] + 1:hi[2] - 1] = True
        vol |= box & ~inner
        vol &= ~inner | (rng.random(vol.shape) < 0.1)
    return vol


def serpentine(nb, width_gap=1):
    """A solid volume carved by one corridor that winds along x through every
    other z row of the y = 1 plane; the path is about V0 * V2 / 2 voxels."""
    X, Y, Z = 4 * nb[0], 4 * nb[1], 4 * nb[2]
    vol = np.ones((X, Y, Z), bool)
    y = 1
    for z in range(1, Z - 1, 2):
        vol[1:X - 1, y, z] = False
        if z + 2 < Z - 1:
            link_x = X - 2 if (z // 2) % 2 == 0 else 1
            vol[link_x, y, z + 1] = False
    return vol
