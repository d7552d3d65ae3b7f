"""GPU parity of the render path: libdroplet_hip.so (through the C ABI) against
the CPU oracle on identical seeded inputs.

Bar: sort keys, pair lists, tile ranges, projected records, the FP32
accumulator and the RGBA8 image are bit-exact (the kernels and the oracle use
one exp specification and IEEE-exact +,-,*,/,sqrt without contraction).
"""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.renderer import Renderer, RendererOptions, Scene, debug_math, scan_u32, sort_pairs_u32
from droplet_amd.types import BackgroundRGBA, SplatData

pytestmark = pytest.mark.gpu


def _gpu_taps(renderer, data, cam, bg=None):
    renderer.set_taps(True)
    try:
        rgba = renderer.render(data, cam, bg or BackgroundRGBA())
        taps = {name: renderer.tap(name) for name in
                ("sorted_indices", "depth_keys", "projected", "tile_counts", "pair_tiles", "pair_splats",
                 "tile_offsets", "accum")}
        st = renderer.stats()
    finally:
        renderer.set_taps(False)
    return rgba, taps, st


def assert_render_parity(renderer, data, cam, bg=None, chunk_cap=None, max_coverage=0):
    rgba, t, st = _gpu_taps(renderer, data, cam, bg)
    ref = oracle.render(data, cam, background=bg, chunk_cap=chunk_cap or 0xFFFFFFFF, max_coverage=max_coverage)
    print(f"case: n={data.count} {cam.width}x{cam.height} visible={ref['num_visible']} pairs={ref['num_pairs']}",
          flush=True)
    assert st["num_visible"] == ref["num_visible"]
    assert st["num_pairs"] == ref["num_pairs"]
    np.testing.assert_array_equal(t["sorted_indices"], ref["sorted_indices"])
    np.testing.assert_array_equal(t["depth_keys"], ref["depth_keys"])
    proj = t["projected"].reshape(-1, 12)
    rp = ref["projected"]
    culled = rp[:, 2] < 0
    # culled splats: only slot 0 is defined (project.ts:166-168); slot 2.w is
    # the reference's padding, which the device uses for the rasterizer's
    # skip threshold
    np.testing.assert_array_equal(proj[culled, :4].view(np.uint32), rp[culled, :4].view(np.uint32))
    np.testing.assert_array_equal(proj[~culled, :11].view(np.uint32), rp[~culled, :11].view(np.uint32))
    np.testing.assert_array_equal(t["tile_counts"], ref["tile_counts"])
    np.testing.assert_array_equal(t["pair_tiles"], ref["pair_tiles"])
    np.testing.assert_array_equal(t["pair_splats"], ref["pair_splats"])
    np.testing.assert_array_equal(t["tile_offsets"], ref["tile_offsets"].reshape(-1))
    np.testing.assert_array_equal(t["accum"].view(np.uint32), ref["accum"].reshape(-1).view(np.uint32))
    np.testing.assert_array_equal(rgba, ref["rgba"])
    return rgba, ref


def test_exp_matches_oracle_bitexact(device):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-110, 95, 200_000), rng.uniform(-5, 0, 200_000),
                        np.array([0.0, -0.0, 1.0, -1.0, -4.5, 88.7, 88.8, -87.4, -103.9, -104.5, 90, -200])]
                       ).astype(np.float32)
    got = debug_math(device, "exp", x)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.expf(x).view(np.uint32))


def test_rasterizer_exp_matches_oracle_bitexact(device):
    """The rasterizer's exp (one-step 2^k, arguments in [-87, 0] or NaN) is
    bit-identical to the exp specification there."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-87, 0, 300_000), rng.uniform(-5, 0, 300_000),
                        -np.abs(rng.standard_normal(100_000)) * 1e-6,
                        np.array([0.0, -0.0, -1e-45, -4.6, -86.99, -87.0])]).astype(np.float32)
    got = debug_math(device, "exp_nonpos", x)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.expf(x).view(np.uint32))
    assert np.isnan(debug_math(device, "exp_nonpos", np.array([np.nan], np.float32))[0])


def test_ieee_primitives_match_host(device):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(300_000) * 10.0 ** rng.uniform(-20, 20, 300_000)).astype(np.float32)
    b = (rng.standard_normal(300_000) * 10.0 ** rng.uniform(-20, 20, 300_000)).astype(np.float32)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(debug_math(device, "div", a, b).view(np.uint32), (a / b).view(np.uint32))
        aa = np.abs(a)
        np.testing.assert_array_equal(debug_math(device, "sqrt", aa).view(np.uint32), np.sqrt(aa).view(np.uint32))


def test_sort_pairs_stable(device):
    rng = np.random.default_rng(2)
    for n, bits in [(1, 32), (17, 4), (5000, 15), (1_000_003, 32), (300_000, 7)]:
        keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        k, v = sort_pairs_u32(device, keys, vals, bits)
        masked = keys & np.uint32((1 << bits) - 1 if bits < 32 else 0xFFFFFFFF)
        order = np.argsort(masked, kind="stable")
        np.testing.assert_array_equal(v, vals[order])
        np.testing.assert_array_equal(k, keys[order])


def test_scan_u32(device):
    """The device-wide scan against a wrapping u32 cumulative sum: empty, one
    tile, tile edges, the largest scan whose tiles sum the earlier tiles'
    totals themselves (4096 tiles), a recursive one (20M values); three
    scans back to back on one workspace."""
    rng = np.random.default_rng(3)
    for n, hi in [(0, 10), (1, 10), (4095, 2 ** 32), (4096, 1000), (4097, 2 ** 32), (64 * 4096 + 5, 2 ** 32),
                  (300_001, 3), (4096 * 4096, 2 ** 32), (20_000_000, 2 ** 32)]:
        print(f"case: scan n={n}", flush=True)
        a = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        out, total = scan_u32(device, a, repeat=3)
        inc = np.cumsum(a, dtype=np.uint32)
        excl = np.concatenate([np.zeros(1, np.uint32), inc[:-1]]) if n else inc
        np.testing.assert_array_equal(out, excl)
        assert total == (int(inc[-1]) if n else 0)


def test_render_plain_workload_parity(renderer):
    data, cam = S.plain_render_workload(seed=0)
    assert_render_parity(renderer, data, cam)


def test_render_small_scenes_parity(renderer):
    cam128 = S.default_camera(128, 128)
    assert_render_parity(renderer, S.one_splat((0, 0, 0)), cam128)
    assert_render_parity(renderer, S.many_splats(5000, "sphere"), S.default_camera(256, 256))
    assert_render_parity(renderer, S.gaussian_cloud(20_000, seed=3, sh_bands=1), S.default_camera(333, 251))
    assert_render_parity(renderer, S.gaussian_cloud(20_000, seed=4, sh_bands=2), S.default_camera(200, 100))
    assert_render_parity(renderer, S.gaussian_cloud(20_000, seed=5, sh_bands=3, pos_std=0.75, pos_clip=3.0),
                         S.default_camera(640, 360), bg=BackgroundRGBA(0.2, 0.4, 0.6, 0.5))


def test_render_equirect_parity(renderer):
    data = S.gaussian_cloud(30_000, seed=6, pos_std=1.5, pos_clip=3.0, sh_bands=1)
    assert_render_parity(renderer, data, S.equirect_camera(512, 256))


def test_one_splat_after_a_large_frame_without_taps(device):
    """Frames without taps after the first run speculatively: the cull writes
    no ids and the depth sort's first pass makes them. A one-splat scene has
    nothing to sort, so the sort must still write its id, or the frame
    would draw the previous frame's splat 0 (regression)."""
    r = Renderer(device)
    try:
        cam = S.default_camera(256, 160)
        big = S.gaussian_cloud(50_000, seed=21)
        np.testing.assert_array_equal(r.render(big, cam), oracle.render(big, cam, taps=False)["rgba"])
        for data in (S.one_splat((0.2, -0.1, 0.0), (0.1, 0.9, 0.3)), S.one_splat((0, 0, 0)),
                     S.gaussian_cloud(2, seed=22), S.empty_scene()):
            got = r.render(data, cam)
            np.testing.assert_array_equal(got, oracle.render(data, cam, taps=False)["rgba"])
    finally:
        r.destroy()


def test_render_empty_scene(renderer):
    rgba = renderer.render(S.empty_scene(), S.default_camera(128, 128))
    img = rgba.reshape(128, 128, 4)
    assert (img[..., :3] <= 5).all()
    assert (img[..., 3] == 255).all()


def test_chunked_matches_unchunked(device):
    data = S.many_splats(200_000, "random")
    cam = S.default_camera(64, 64)
    single = Renderer(device)
    multi = Renderer(device, RendererOptions(chunk_cap=50_000))
    try:
        a = single.render(data, cam)
        b = multi.render(data, cam)
        np.testing.assert_array_equal(a, b)
        assert multi.stats()["num_chunks"] == 4
        ref = oracle.render(data, cam, chunk_cap=50_000, taps=False)
        np.testing.assert_array_equal(b, ref["rgba"])
    finally:
        single.destroy()
        multi.destroy()


def test_chunked_taps_match_oracle(device):
    data = S.gaussian_cloud(30_000, seed=8)
    cam = S.default_camera(320, 180)
    r = Renderer(device, RendererOptions(chunk_cap=7_000))
    try:
        rgba, t, st = _gpu_taps(r, data, cam)
        ref = oracle.render(data, cam, chunk_cap=7_000)
        assert st["num_chunks"] == ref["num_chunks"] == 5
        np.testing.assert_array_equal(t["pair_tiles"], ref["pair_tiles"])
        np.testing.assert_array_equal(t["pair_splats"], ref["pair_splats"])
        np.testing.assert_array_equal(t["tile_offsets"], ref["tile_offsets"].reshape(-1))
        np.testing.assert_array_equal(rgba, ref["rgba"])
    finally:
        r.destroy()


def test_scene_handle_matches_render(device, renderer):
    data, cam = S.plain_render_workload(seed=11)
    scene = Scene(device, data)
    try:
        a = renderer.render_scene(scene, cam)
        b = renderer.render(data, cam)
        np.testing.assert_array_equal(a, b)
        renderer.render_scene(scene, cam, readback=False)
        np.testing.assert_array_equal(renderer.read_image(), a)
    finally:
        scene.destroy()


def test_max_coverage_truncation_parity(device):
    data = S.dense_cluster(20_000)
    cam = S.default_camera(256, 256)
    r = Renderer(device, RendererOptions(max_coverage_per_splat=7))
    try:
        assert_render_parity(r, data, cam, max_coverage=7)
    finally:
        r.destroy()


@pytest.mark.parametrize("binning", ["span", "sort"])
def test_binning_paths_match_oracle(device, binning):
    r = Renderer(device, RendererOptions(binning=binning))
    try:
        data, cam = S.plain_render_workload(seed=12)
        assert_render_parity(r, data, cam)
        assert_render_parity(r, S.dense_cluster(20_000), S.default_camera(256, 256))
        # maxCoverage truncation leaves a partial last row of tiles
        r2 = Renderer(device, RendererOptions(binning=binning, max_coverage_per_splat=5))
        try:
            assert_render_parity(r2, S.gaussian_cloud(5000, seed=13, log_scale_range=(-4.0, -2.0)),
                                 S.default_camera(300, 200), max_coverage=5)
        finally:
            r2.destroy()
    finally:
        r.destroy()


def test_wide_grid_falls_back_to_sort(renderer):
    # 4400 px = 275 tile columns: above the span path's 256-column LDS limit
    data = S.gaussian_cloud(20_000, seed=14)
    assert_render_parity(renderer, data, S.default_camera(4400, 64))


def test_rasterize_work_count_matches_oracle(device):
    """The instrumented kernel counts the reference's per-pixel loop
    iterations (executed before the T < 1e-4 break); strip culling and the
    exp skip inside the kernel must not change that count."""
    data, cam = S.plain_render_workload(seed=0)
    r = Renderer(device)
    try:
        r.set_counting(True)
        img = r.render(data, cam)
        st = r.stats()
        iters = st["pixel_iters"]
    finally:
        r.destroy()
    ref = oracle.render(data, cam, taps=False)
    assert iters == ref["pixel_iters"]
    assert 0 < st["pairs_consumed"] <= st["num_pairs"]
    np.testing.assert_array_equal(img, ref["rgba"])


def test_exp_skip_threshold_is_conservative(renderer):
    """The device stores a per-splat power threshold tau in record slot 2.w and
    skips the pixel body for power < tau. Check on every projected splat that
    at the largest f32 power below tau the reference's alpha,
    min(0.999, a*max(0, exp(power) - floor)), is already below 1/255 (alpha is
    monotone in power, so no skipped pixel could have composited)."""
    data, cam = S.plain_render_workload(seed=15)
    _, t, _ = _gpu_taps(renderer, data, cam)
    rec = t["projected"].reshape(-1, 12)
    rec = rec[rec[:, 2] > 0]
    a = rec[:, 7].astype(np.float32)
    tau = rec[:, 11].astype(np.float32)
    finite = np.isfinite(tau)
    assert finite.mean() > 0.9
    below = np.nextafter(tau[finite], np.float32(-np.inf)).astype(np.float32)
    below = np.minimum(below, np.float32(0.0))
    e = oracle.expf(below)
    f32 = np.float32
    alpha = np.minimum(f32(0.999), a[finite] * np.maximum(f32(0.0), e - f32(1.1108996538e-2)))
    assert (alpha < f32(0.00392156862745098)).all()
    # and the threshold is tight: just above tau, most splats do composite
    above = np.nextafter(tau[finite], np.float32(np.inf)).astype(np.float32) + np.float32(1e-3)
    above = np.minimum(above, np.float32(0.0))
    e2 = oracle.expf(above)
    alpha2 = np.minimum(f32(0.999), a[finite] * np.maximum(f32(0.0), e2 - f32(1.1108996538e-2)))
    assert (alpha2 >= f32(0.00392156862745098)).mean() > 0.95
    # tau is +inf only where alpha == 0
    assert (a[~finite] == 0).all()


@pytest.mark.parametrize("first_pass,growth", [(None, None), (1000, 2), (5000, 1), (50_000, 8)])
def test_front_to_back_passes_match_oracle(device, first_pass, growth):
    """Span binning in front-to-back passes closes tiles whose pixels have all
    saturated and bins nothing into them afterwards. The image and the
    reference's per-pixel loop count are unchanged."""
    r = Renderer(device, RendererOptions(first_pass=first_pass, pass_growth=growth))
    try:
        for data, cam in [S.plain_render_workload(seed=21), (S.dense_cluster(30_000), S.default_camera(256, 256)),
                          (S.gaussian_cloud(60_000, seed=22, sh_bands=2, pos_std=0.75, pos_clip=3.0),
                           S.default_camera(640, 360))]:
            r.set_counting(True)
            img = r.render(data, cam)
            st = r.stats()
            r.set_counting(False)
            ref = oracle.render(data, cam, taps=False)
            print(f"case: n={data.count} {cam.width}x{cam.height} passes={st['num_chunks']} "
                  f"entries={st['list_entries']} of {st['num_pairs']}", flush=True)
            np.testing.assert_array_equal(img, ref["rgba"])
            assert st["pixel_iters"] == ref["pixel_iters"]
            assert st["num_pairs"] == ref["num_pairs"]
            assert st["list_entries"] <= st["num_pairs"]
            if first_pass is not None and st["num_visible"] > first_pass:
                assert st["num_chunks"] > 1
    finally:
        r.destroy()


@pytest.mark.parametrize("bands", [0, 1, 3])
def test_repeated_frames_match_oracle(device, bands):
    """The first frame of a renderer reads the visible and pair totals after
    the depth sort; later frames of the same size do not (capacity from the
    earlier frames, checked at the end) and, in passes, compute colours only
    for the splats that reach a tile list. Every frame equals the oracle, and
    a scene with more pairs than the capacity is rendered again exactly."""
    data = S.gaussian_cloud(80_000, seed=70 + bands, sh_bands=bands, pos_std=0.75, pos_clip=3.0)
    cam = S.default_camera(640, 360)
    big = S.gaussian_cloud(160_000, seed=80 + bands, sh_bands=bands, pos_std=0.75, pos_clip=3.0)
    r = Renderer(device, RendererOptions(first_pass=8_000, pass_growth=4))
    try:
        ref = oracle.render(data, cam, taps=False)
        for frame in range(3):
            img = r.render(data, cam)
            st = r.stats()
            print(f"case: bands={bands} frame={frame} passes={st['num_chunks']}", flush=True)
            np.testing.assert_array_equal(img, ref["rgba"])
            assert st["num_pairs"] == ref["num_pairs"]
            assert st["num_chunks"] > 1
        img = r.render(big, cam)
        np.testing.assert_array_equal(img, oracle.render(big, cam, taps=False)["rgba"])
        assert r.stats()["num_pairs"] > ref["num_pairs"]
    finally:
        r.destroy()


def test_depth_key_span_speculation_matches_oracle(device):
    """After a frame, the renderer sorts on depth keys less their minimum over
    the last frame's key span (three 9-bit passes instead of four 8-bit
    ones). The order equals the 32-bit sort's while the span fits; a frame
    whose keys span more bits is rendered again with the full sort. Splats
    behind the camera (invisible) must still sort after every visible one."""
    cam = S.default_camera(320, 180)
    # a narrow depth range (22-bit key span), with every 50th splat moved
    # behind the camera (z = 6 > 4): invisible splats inside a short span
    narrow = S.gaussian_cloud(60_000, seed=91, sh_bands=1, pos_std=0.15, pos_clip=0.4)
    pos = narrow.positions.reshape(-1, 3).copy()
    pos[::50, 2] = 6.0
    narrow = SplatData(narrow.count, pos.reshape(-1), narrow.rotations, narrow.log_scales, narrow.opacity_logits,
                       narrow.colors_dc, narrow.colors_sh, narrow.sh_bands)
    wide = S.gaussian_cloud(60_000, seed=92, sh_bands=1, pos_std=3.0, pos_clip=9.0)  # 27-bit span
    r = Renderer(device)
    try:
        for name, data, frames in (("narrow", narrow, 3), ("wide", wide, 3), ("narrow", narrow, 2)):
            ref = oracle.render(data, cam, taps=False)
            assert ref["num_visible"] < data.count
            for frame in range(frames):
                img = r.render(data, cam)
                st = r.stats()
                print(f"case: span {name} frame={frame} visible={st['num_visible']}", flush=True)
                np.testing.assert_array_equal(img, ref["rgba"])
                assert st["num_visible"] == ref["num_visible"]
    finally:
        r.destroy()


def test_front_to_back_passes_skip_closed_tiles(device):
    """On a dense scene many (tile, splat) pairs lie behind saturated pixels:
    the passes build fewer list entries than the full lists hold (tiles at
    the cluster's rim never saturate and keep their whole lists)."""
    data = S.dense_cluster(200_000)
    cam = S.default_camera(512, 512)
    r = Renderer(device, RendererOptions(first_pass=20_000, pass_growth=4))
    full = Renderer(device, RendererOptions(full_lists=True))
    try:
        a = r.render(data, cam)
        st = r.stats()
        b = full.render(data, cam)
        fst = full.stats()
        np.testing.assert_array_equal(a, b)
        assert fst["list_entries"] == fst["num_pairs"] == st["num_pairs"]
        assert st["list_entries"] < 0.9 * st["num_pairs"]
    finally:
        r.destroy()
        full.destroy()


def test_autotune_keeps_the_image(device):
    data = S.gaussian_cloud(150_000, seed=41, pos_std=0.6, pos_clip=2.5)
    cam = S.default_camera(640, 360)
    scene = Scene(device, data)
    r = Renderer(device)
    try:
        before = r.render_scene(scene, cam)
        res = r.autotune(scene, cam)
        print(f"case: autotune {res}", flush=True)
        assert res["candidates"] >= 10
        assert res["chosen_ms"] <= res["default_ms"]
        after = r.render_scene(scene, cam)
        np.testing.assert_array_equal(after, before)
        np.testing.assert_array_equal(after, oracle.render(data, cam, taps=False)["rgba"])
    finally:
        r.destroy()
        scene.destroy()


@pytest.mark.parametrize("digit_bits", [4, 5, 7])
def test_sort_digit_width_keeps_the_image(device, digit_bits):
    data, cam = S.plain_render_workload(seed=42)
    r = Renderer(device, RendererOptions(sort_digit_bits=digit_bits))
    try:
        assert_render_parity(r, data, cam)
    finally:
        r.destroy()
