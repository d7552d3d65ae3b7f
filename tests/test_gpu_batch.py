"""GPU parity of the batched views (dl_batch_*): every view's image equals
the single-view renderer's image of the same (scene, camera, background),
bit for bit, directly and on graph replays with new cameras."""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.batch import BatchRenderer, View
from droplet_amd.math import mat4_look_at
from droplet_amd.renderer import Renderer, RendererOptions, Scene
from droplet_amd.types import BackgroundRGBA, RenderCamera
from droplet_amd._lib import DropletError

pytestmark = pytest.mark.gpu

SIZES = ((160, 90), (320, 180), (200, 200))


def _orbit(k, count, radius=4.0, phase=0.0, size=None):
    a = 2.0 * math.pi * k / count + phase
    eye = (radius * math.sin(a), 0.3 * radius * math.sin(2 * a), radius * math.cos(a))
    w, h = size or SIZES[k % len(SIZES)]
    return RenderCamera(w, h, mat4_look_at(eye, (0, 0, 0), (0, 1, 0)), math.pi / 3, 0.01, 1000.0)


@pytest.fixture(scope="module")
def scenes(device):
    datas = [S.gaussian_cloud(20_000, seed=31), S.gaussian_cloud(3_000, seed=32, sh_bands=1),
             S.empty_scene(), S.gaussian_cloud(50_000, seed=33, sh_bands=3, pos_std=0.75, pos_clip=3.0)]
    sc = [Scene(device, d) for d in datas]
    yield datas, sc
    for s in sc:
        s.destroy()


def _views(phase=0.0, cams=6):
    views = []
    for si in range(4):
        for k in range(cams):
            bg = BackgroundRGBA(0.1 * k, 0.2, 0.3, 1.0) if k % 2 else None
            views.append(View(si, _orbit(k, cams, phase=phase + 0.1 * si), bg))
    return views


def _check_against_plain(device, datas, batch, views):
    r = Renderer(device)
    try:
        for k, v in enumerate(views):
            ref = r.render(datas[v.scene], v.camera, v.background or BackgroundRGBA())
            got = batch.image(k)
            print(f"case: view {k} scene {v.scene} {v.camera.width}x{v.camera.height}", flush=True)
            np.testing.assert_array_equal(got, ref)
    finally:
        r.destroy()


def test_batch_matches_single_view_renders(device, scenes):
    datas, sc = scenes
    b = BatchRenderer(device, use_graph=False)
    try:
        views = _views()
        b.render(sc, views)
        _check_against_plain(device, datas, b, views)
        # one view also against the oracle
        v = views[1]
        ref = oracle.render(datas[v.scene], v.camera, background=v.background, taps=False)
        np.testing.assert_array_equal(b.image(1), ref["rgba"])
        st = b.stats()
        assert st["num_views"] == len(views) and st["captures"] == 0
    finally:
        b.destroy()


def test_batch_graph_replay_with_new_cameras(device, scenes):
    datas, sc = scenes
    b = BatchRenderer(device, use_graph=True)
    try:
        v1 = _views(0.0)
        b.render(sc, v1)
        _check_against_plain(device, datas, b, v1)
        assert b.stats()["captures"] == 1
        for rep in range(3):
            v2 = _views(0.37 * (rep + 1))
            b.render(sc, v2)
            st = b.stats()
            assert st["graph_replays"] == rep + 1
            _check_against_plain(device, datas, b, v2)
    finally:
        b.destroy()


def test_batch_regrows_lists_when_cameras_move_closer(device, scenes):
    datas, sc = scenes
    b = BatchRenderer(device, use_graph=True)
    try:
        far = [View(0, _orbit(k, 4, radius=12.0, size=(320, 180))) for k in range(4)]
        b.render(sc, far)
        near = [View(0, _orbit(k, 4, radius=2.5, size=(320, 180))) for k in range(4)]
        b.render(sc, near)
        _check_against_plain(device, datas, b, near)
        b.render(sc, near)  # re-captured plan replays
        assert b.stats()["graph_replays"] >= 1
        _check_against_plain(device, datas, b, near)
    finally:
        b.destroy()


def test_batch_groups(device, scenes):
    """Small groups (the option caps scene splats per group) give the same
    images."""
    datas, sc = scenes
    b = BatchRenderer(device, RendererOptions(chunk_cap=25_000), use_graph=True)
    try:
        views = _views(0.2, cams=3)
        b.render(sc, views)
        assert b.stats()["num_groups"] > 2
        _check_against_plain(device, datas, b, views)
        b.render(sc, views)
        _check_against_plain(device, datas, b, views)
    finally:
        b.destroy()


def test_batch_equirect_and_coverage_cap(device, scenes):
    datas, sc = scenes
    b = BatchRenderer(device, RendererOptions(max_coverage_per_splat=6), use_graph=False)
    r = Renderer(device, RendererOptions(max_coverage_per_splat=6))
    try:
        views = [View(3, S.equirect_camera(512, 256)), View(0, _orbit(1, 5, radius=2.0, size=(300, 200)))]
        b.render(sc, views)
        for k, v in enumerate(views):
            np.testing.assert_array_equal(b.image(k), r.render(datas[v.scene], v.camera))
    finally:
        b.destroy()
        r.destroy()


def test_batch_rejects_oversized_view(device, scenes):
    _, sc = scenes
    b = BatchRenderer(device)
    try:
        with pytest.raises(DropletError):
            b.render(sc, [View(0, _orbit(0, 1, size=(5000, 100)))])
        with pytest.raises(DropletError):
            b.render(sc, [View(7, _orbit(0, 1))])
    finally:
        b.destroy()


def test_batch_autotune_keeps_images(device, scenes):
    datas, sc = scenes
    b = BatchRenderer(device)
    try:
        views = _views(0.5, cams=3)
        res = b.autotune(sc, views)
        print(f"case: batch autotune {res}", flush=True)
        assert res["candidates"] >= 5
        b.render(sc, views)
        _check_against_plain(device, datas, b, views)
    finally:
        b.destroy()


def test_batch_splits_groups_over_the_pair_budget(device, scenes):
    """Groups above the pair budget split by views; a view above it alone
    renders through the single-view path (which cuts its own passes by the
    same budget) into its image slot. A budget at the median view's pair
    count forces both; every view still equals the single-view render, on
    the direct call, the graph replay and a call with new cameras."""
    datas, sc = scenes
    views = _views()
    pairs = [oracle.render(datas[v.scene], v.camera, taps=False)["num_pairs"] for v in views]
    budget = int(sorted(p for p in pairs if p)[len([p for p in pairs if p]) // 2])
    print(f"case: budget {budget} pairs; views over it: {sum(p > budget for p in pairs)} of {len(views)}", flush=True)
    assert sum(p > budget for p in pairs) >= 3
    b = BatchRenderer(device, use_graph=True)
    try:
        b.set_pair_budget(budget)
        for call, vs in enumerate((views, views, _views(phase=0.4))):
            b.render(sc, vs)
            st = b.stats()
            print(f"case: call {call}: {st['num_groups']} groups, {st['graph_replays']} replays", flush=True)
            assert st["num_groups"] >= 2
            _check_against_plain(device, datas, b, vs)
            assert st["num_pairs"] == sum(oracle.render(datas[v.scene], v.camera, taps=False)["num_pairs"]
                                          for v in vs)
    finally:
        b.destroy()
