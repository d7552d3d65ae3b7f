"""Tile ranges at the autotuner's 8- and 32-px tiles, on the CPU.

The device lists a splat at 8 px in the 8-px tiles of its 16-px range
narrowed to floor((s -+ r) / 8) when s -+ r and r lie below 2^22 in magnitude,
and at 32 px in the 32-px tiles holding a tile of its 16-px range (the
16-px range is the reference's, src/render/shaders/project.ts:223-233). The
8-px narrowing is safe only if every pixel it drops fails the reference's
box test |px + 0.5 - sx| > r in f32 (rasterize.ts:66-69). This file restates
the range in numpy float32 and checks that claim by brute force on splats
placed to make s - r and s + r land on, and one ulp beside, the 8-px tile
edges, at magnitudes up to 2^23. It also checks that the tuner's candidate
list (dl_renderer_tune_candidates) holds every tile size and workgroup.
"""
import os

import numpy as np
import pytest

from droplet_amd import _lib as L

F = np.float32
LIM = F(4194304.0)


def range16(s, r, t16):
    lo = np.maximum(F(0), np.floor((s - r) / F(16)))
    hi = np.minimum(F(t16 - 1), np.floor((s + r) / F(16)))
    return lo, hi


def range8(s, r, t8, t16):
    a0, a1 = range16(s, r, t16)
    x0, x1 = 2 * a0, np.minimum(2 * a1 + 1, F(t8 - 1))
    e, f = s - r, s + r
    ok = (np.abs(e) < LIM) & (np.abs(f) < LIM) & (r < LIM)
    n0 = np.maximum(F(0), np.floor(e / F(8)))
    n1 = np.maximum(F(0), np.floor(f / F(8)))
    x0 = np.where(ok, np.maximum(x0, n0), x0)
    x1 = np.where(ok, np.minimum(x1, n1), x1)
    return a0, a1, x0, x1


def adversarial_splats(rng, n):
    """s -+ r on an 8-px edge, or an ulp to either side, for radii from
    1/4 px to 2^22 and edges from -2^23 to 2^23."""
    k = rng.integers(-(1 << 20), 1 << 20, n).astype(np.float64)
    scale = 2.0 ** rng.integers(-20, 3, n)
    edge = (k * scale * 8.0).astype(F)
    r = (2.0 ** rng.uniform(-2, 22, n)).astype(F)
    r = np.ceil(r)  # the projection's radius is ceil(...)
    side = rng.integers(0, 2, n)
    s = np.where(side == 0, edge + r, edge - r).astype(F)  # s - r or s + r near the edge
    nudge = rng.integers(-1, 2, n)
    s = np.where(nudge < 0, np.nextafter(s, F(-np.inf)), np.where(nudge > 0, np.nextafter(s, F(np.inf)), s))
    return s.astype(F), r.astype(F)


@pytest.mark.parametrize("width", [2048, 2049, 8191])
def test_eight_px_narrowing_drops_only_rejected_pixels(width):
    rng = np.random.default_rng(width)
    s, r = adversarial_splats(rng, 20000)
    # also splats centred on the image with random subpixel positions
    s2 = rng.uniform(-64, width + 64, 20000).astype(F)
    r2 = np.ceil(rng.uniform(0.2, 40, 20000)).astype(F)
    s, r = np.concatenate([s, s2]), np.concatenate([r, r2])
    t16, t8 = (width + 15) // 16, (width + 7) // 8
    a0, a1, x0, x1 = range8(s, r, t8, t16)
    nonempty = a1 >= a0
    assert (x0[nonempty] >= 2 * a0[nonempty]).all() and (x1[nonempty] <= 2 * a1[nonempty] + 1).all()
    checked = 0
    for j in np.nonzero(nonempty)[0]:
        lo16, hi16 = int(a0[j]) * 16, min(int(a1[j]) * 16 + 15, width - 1)
        lo8, hi8 = int(x0[j]) * 8, int(x1[j]) * 8 + 7
        # pixels of the 16-px range that the 8-px range drops: at most 15
        # on each side (or all of them when the narrowed range is empty)
        if x1[j] < x0[j]:
            drop = np.arange(lo16, hi16 + 1)
        else:
            drop = np.concatenate([np.arange(lo16, min(lo8, hi16 + 1)), np.arange(max(hi8 + 1, lo16), hi16 + 1)])
        if drop.size == 0:
            continue
        dx = (drop.astype(F) + F(0.5)) - s[j]
        assert (np.abs(dx) > r[j]).all(), (j, s[j], r[j], drop[np.abs(dx) <= r[j]][:4])
        checked += drop.size
    assert checked > 100_000


def test_thirty_two_px_range_halves_the_sixteen_px_range():
    rng = np.random.default_rng(5)
    s = rng.uniform(-100, 4200, 50000).astype(F)
    r = np.ceil(rng.uniform(0.2, 300, 50000)).astype(F)
    a0, a1 = range16(s, r, (4096 + 15) // 16)
    ok = a1 >= a0
    # every 16-px tile of the range lies in a listed 32-px tile, and the
    # 32-px range never holds more tiles
    b0, b1 = np.floor(a0 / 2), np.floor(a1 / 2)
    assert ((b1 - b0 + 1)[ok] <= (a1 - a0 + 1)[ok]).all()
    assert (b1[ok] <= (4096 + 31) // 32 - 1).all()


def test_tune_candidates_hold_every_shape_and_digit_width():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libdroplet_hip.so is not built")
    from droplet_amd.renderer import tune_candidates
    cands = tune_candidates(None, 6_000_000)
    assert {c.sort_digit_bits or 8 for c in cands} == {4, 6, 8}
    shapes = {(c.tile_size or 16, c.raster_group or 1) for c in cands if c.tile_size}
    assert shapes == {(16, 1), (16, 2), (8, 1), (8, 2), (32, 1)} - {(16, 1)}
    # stage 1 keeps the base's shape, stage 2 the base's schedule
    assert all(not c.tile_size for c in cands[:-len(shapes)])
    assert all(c.first_pass is None and not c.full_lists for c in cands[-len(shapes):])
    assert any(c.full_lists for c in cands) and any(c.first_pass == 6_000_000 // 16 for c in cands)
