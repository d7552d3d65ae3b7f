"""Tile sizes 8 and 32 px and 2x2-tile workgroups (the autotuner's shape
candidates) against the oracle, which rasterizes 16-px tiles as the
reference does (src/render/shaders/constants.ts:19, rasterize.ts:36).

The other sizes change the tile lists but not the image: without maxCoverage
truncation every size lists a splat in each tile holding a pixel that the
reference blends it into, and the 32-px rasterizer keeps each 16-px quadrant
to the splats whose 16-px range holds it. A frame in which maxCoverage
truncates a splat at 16-px tiles renders at 16 px. Bar: RGBA8 bit-exact.
"""
import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.renderer import Renderer, RendererOptions, Scene, tune_candidates

pytestmark = pytest.mark.gpu

SHAPES = [(8, 1), (8, 2), (16, 2), (32, 1)]


def _cases():
    """(name, data, camera, extra options, frames). The 3840x2160 case runs
    passes with closed tiles and live-splat compaction over three frames
    (speculative capacity, key-span sort, work-ordered launch: at least 4096
    launch units at every shape); 4400 px is past the span path's 256 tile
    columns at 16 and 8 px but not at 32 px."""
    return [
        ("plain", *S.plain_render_workload(seed=31), {}, 1),
        ("dense passes", S.dense_cluster(30_000), S.default_camera(256, 256),
         {"first_pass": 5000, "pass_growth": 2}, 2),
        ("4k passes", S.gaussian_cloud(200_000, seed=32, sh_bands=1, pos_std=0.75, pos_clip=3.0),
         S.default_camera(3840, 2160), {"first_pass": 20_000, "pass_growth": 4}, 3),
        ("equirect", S.gaussian_cloud(30_000, seed=33, pos_std=1.5, pos_clip=3.0, sh_bands=1),
         S.equirect_camera(1024, 512), {}, 2),
        ("sort binning", S.gaussian_cloud(20_000, seed=34), S.default_camera(640, 360), {"binning": "sort"}, 1),
        ("wide", S.gaussian_cloud(20_000, seed=35), S.default_camera(4400, 64), {}, 1),
    ]


@pytest.mark.parametrize("case", range(6))
def test_tile_shapes_keep_the_image(device, case):
    name, data, cam, extra, frames = _cases()[case]
    ref = oracle.render(data, cam, taps=False)["rgba"]
    scene = Scene(device, data)
    try:
        for tile, group in SHAPES:
            r = Renderer(device, RendererOptions(tile_size=tile, raster_group=group, **extra))
            try:
                for frame in range(frames):
                    img = r.render_scene(scene, cam)
                    st = r.stats()
                    print(f"case: {name} tile={tile} group={group} frame={frame} passes={st['num_chunks']} "
                          f"pairs={st['num_pairs']} tiles={st['tiles_x']}x{st['tiles_y']}", flush=True)
                    assert st["tile_size"] == tile and not st["coverage_truncated"]
                    assert st["raster_group"] == (1 if tile == 32 else group)
                    np.testing.assert_array_equal(img, ref)
            finally:
                r.destroy()
    finally:
        scene.destroy()


def test_tile_shapes_list_the_expected_pairs(device):
    """Pair totals at 8 and 32 px follow from the 16-px ranges: a 32-px range
    is the 16-px range halved, so it never holds more tiles; an 8-px range
    lies inside the doubled 16-px range, so it holds at most four times as
    many."""
    data, cam = S.plain_render_workload(seed=36)
    pairs = {}
    for tile in (8, 16, 32):
        r = Renderer(device, RendererOptions(tile_size=tile, full_lists=True))
        try:
            r.render(data, cam)
            pairs[tile] = r.stats()["num_pairs"]
        finally:
            r.destroy()
    ref = oracle.render(data, cam, taps=False)
    print(f"case: pairs by tile size {pairs}", flush=True)
    assert pairs[16] == ref["num_pairs"]
    assert pairs[32] <= pairs[16] <= pairs[8] <= 4 * pairs[16]


def test_tile_shape_falls_back_where_max_coverage_truncates(device):
    """maxCoverage 7 truncates 2015 splats of the dense cluster at 16-px
    tiles (none of the clean cloud): those frames render at 16 px (a
    renderer's first frame finds out after its cull, a later speculative one
    at its end), the frame after them starts at 16 px, and the next one
    returns to the asked size."""
    cam = S.default_camera(256, 256)
    trunc = S.dense_cluster(20_000)
    clean = S.gaussian_cloud(5000, seed=37, log_scale_range=(-6.0, -5.0))
    refs = {"trunc": oracle.render(trunc, cam, taps=False, max_coverage=7)["rgba"],
            "clean": oracle.render(clean, cam, taps=False, max_coverage=7)["rgba"]}
    for tile in (8, 32):
        r = Renderer(device, RendererOptions(tile_size=tile, max_coverage_per_splat=7))
        try:
            seq = [("trunc", 16, True), ("clean", 16, False), ("clean", tile, False), ("clean", tile, False),
                   ("trunc", 16, True), ("trunc", 16, True), ("clean", 16, False), ("clean", tile, False)]
            for k, (which, want_tile, truncated) in enumerate(seq):
                img = r.render(trunc if which == "trunc" else clean, cam)
                st = r.stats()
                print(f"case: tile={tile} frame={k} {which} -> tile {st['tile_size']} "
                      f"truncated={st['coverage_truncated']}", flush=True)
                assert st["tile_size"] == want_tile
                assert st["coverage_truncated"] == truncated
                np.testing.assert_array_equal(img, refs[which])
        finally:
            r.destroy()


def test_every_tune_candidate_keeps_the_image(device):
    """Each option set the autotuner times (tune_candidates: digit widths and
    pass schedules, then tile sizes and workgroups) renders the oracle's
    image, and so does every shape with a multi-pass schedule."""
    data = S.gaussian_cloud(150_000, seed=41, pos_std=0.6, pos_clip=2.5)
    cam = S.default_camera(640, 360)
    ref = oracle.render(data, cam, taps=False)["rgba"]
    cands = tune_candidates(None, data.count)
    digits = {c.sort_digit_bits or 8 for c in cands}
    shapes = {(c.tile_size or 16, c.raster_group or 1) for c in cands}
    assert digits == {4, 6, 8} and {(8, 1), (8, 2), (16, 2), (32, 1)} <= shapes
    extra = [RendererOptions(tile_size=t, raster_group=g, first_pass=20_000, pass_growth=4, sort_digit_bits=6)
             for t, g in SHAPES]
    scene = Scene(device, data)
    try:
        for k, opts in enumerate(cands + extra):
            r = Renderer(device, opts)
            try:
                for frame in range(2):
                    img = r.render_scene(scene, cam)
                    print(f"case: candidate {k} {opts} frame={frame}", flush=True)
                    np.testing.assert_array_equal(img, ref)
            finally:
                r.destroy()
    finally:
        scene.destroy()
