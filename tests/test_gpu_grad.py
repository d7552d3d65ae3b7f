"""GPU checks of the differentiable render (dl_render_forward /
dl_render_backward): the forward image equals the oracle's pre-quantisation
image bit for bit. The backward gradients match FP64 central differences of
the same forward (oracle/gradcheck.c, h = 1e-4) on random parameters of
splats that reach the oracle's tile lists. A parameter whose stencil crosses
a threshold (integer radius, depth order, culling, a per-pixel branch where
the f32 and f64 forwards differ) is differenced again at h = 1e-5 and 1e-6;
one still flagged after that is left out, and at least 90% of the sampled
parameters must be checked."""
import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.renderer import Renderer, Scene
from droplet_amd.types import SplatData

pytestmark = pytest.mark.gpu

KINDS = (("pos", 0, 3), ("rot", 3, 4), ("logscale", 7, 3), ("logit", 10, 1), ("dc", 11, 3), ("sh", 14, None))
REL_TOL = 1e-3


def _flat(grads, n, m):
    """(n, 14 + 3m) parameter-gradient table in the checker's order."""
    cols = [grads["positions"].reshape(n, 3), grads["rotations"].reshape(n, 4), grads["log_scales"].reshape(n, 3),
            grads["opacity_logits"].reshape(n, 1), grads["colors_dc"].reshape(n, 3)]
    if m:
        cols.append(grads["colors_sh"].reshape(n, 3 * m))
    return np.concatenate(cols, axis=1).astype(np.float64)


def _perturbed(data, seed):
    rng = np.random.default_rng(seed)
    pos = (data.positions + rng.normal(0.0, 0.01, data.positions.size)).astype(np.float32)
    op = (data.opacity_logits + rng.normal(0.0, 0.1, data.count)).astype(np.float32)
    return SplatData(data.count, pos, data.rotations, data.log_scales, op, data.colors_dc, data.colors_sh,
                     data.sh_bands)


def _check(device, data, cam, nparams, seed):
    m = {0: 0, 1: 3, 2: 8, 3: 15}[data.sh_bands]
    r = Renderer(device)
    sc = Scene(device, data)
    st = Scene(device, _perturbed(data, seed))
    try:
        target = r.forward(st, cam).copy()
        img = r.forward(sc, cam)
        ref = oracle.render(data, cam, taps=True, chunk_cap=0xFFFFFFFF)
        acc = ref["accum"].reshape(cam.height, cam.width, 4)
        np.testing.assert_array_equal(img, np.clip(acc[..., :3], np.float32(0), np.float32(1)))
        grads = r.backward(2.0 * (img - target))
    finally:
        sc.destroy()
        st.destroy()
        r.destroy()
    table = _flat(grads, data.count, m)
    gc = oracle.GradCheck(data, cam, target)
    # splats that reach a tile list of the oracle's render (independent of
    # the GPU's output, so a splat whose gradient the backward drops entirely
    # is sampled like any other and fails against its non-zero FD value)
    listed = np.unique(ref["sorted_indices"][ref["pair_splats"]])
    rng = np.random.default_rng(seed + 1)
    ids = rng.choice(listed, size=nparams, replace=listed.size < nparams)
    params = rng.integers(0, 14 + 3 * m, size=nparams)
    fd, flags = gc.fd(ids, params, h=1e-4)
    first_flags = {f: int(((flags & f) != 0).sum()) for f in (1, 2, 4, 8, 16, 32)}
    # a stencil that crosses a threshold (radius, depth order, culling, a
    # per-pixel branch, a projection kink) is differenced again with a
    # smaller h, which crosses it with proportionally smaller probability
    for h in (1e-5, 1e-6):
        redo = np.flatnonzero(flags != 0)
        if redo.size == 0:
            break
        fd2, fl2 = gc.fd(ids[redo], params[redo], h=h)
        fd[redo] = fd2
        flags[redo] = fl2
    g = table[ids, params]
    clean = flags == 0
    # relative error with a floor of 1e-3 of the largest |FD| of the same kind,
    # so that a parameter whose true gradient is tiny is judged in absolute terms
    rel = np.empty(nparams)
    for name, lo, cnt in KINDS:
        sel = (params >= lo) & (params < lo + cnt if cnt else params >= lo)
        if not sel.any():
            continue
        floor = 1e-3 * np.abs(fd[sel & clean]).max() if (sel & clean).any() else 1e-30
        rel[sel] = np.abs(g[sel] - fd[sel]) / np.maximum(np.abs(fd[sel]), max(floor, 1e-30))
    worst = np.argsort(-np.where(clean, rel, -1))[:5]
    kinds = {f: int(((flags & f) != 0).sum()) for f in (1, 2, 4, 8, 16, 32)}
    print(f"case: n={data.count} {cam.width}x{cam.height} params={nparams} listed splats={listed.size} "
          f"flagged at h=1e-4: {first_flags}, still flagged after h=1e-5, 1e-6: {int((~clean).sum())} {kinds} "
          f"checked={int(clean.sum())} zero-gradient params={int((g[clean] == 0).sum())} "
          f"rel<=1e-3: {(rel[clean] <= REL_TOL).mean():.4f} max={rel[clean].max():.2e}", flush=True)
    for w in worst:
        print(f"  worst id={ids[w]} param={params[w]} gpu={g[w]:.6e} fd={fd[w]:.6e} rel={rel[w]:.2e}", flush=True)
    return rel[clean], flags


def test_forward_backward_small(device):
    data = S.gaussian_cloud(4000, seed=60, sh_bands=1)
    cam = S.default_camera(192, 128)
    rel, flags = _check(device, data, cam, 600, seed=61)
    assert rel.size >= 540
    assert (rel <= REL_TOL).all()


def test_forward_backward_bands3(device):
    data = S.gaussian_cloud(3000, seed=62, sh_bands=3, pos_std=0.6, pos_clip=2.0)
    cam = S.default_camera(160, 120)
    rel, _ = _check(device, data, cam, 500, seed=63)
    assert rel.size >= 450
    assert (rel <= REL_TOL).all()


def test_forward_backward_backward_workload(device):
    """The backward workload: 1M Gaussians, shBands 1, 1920x1080; target from a
    copy with positions + N(0, 0.01^2) and logits + N(0, 0.1^2); 1000 random
    parameters."""
    data = S.gaussian_cloud(1_000_000, seed=64, sh_bands=1)
    cam = S.default_camera(1920, 1080)
    rel, _ = _check(device, data, cam, 1000, seed=65)
    assert rel.size >= 900
    assert (rel <= REL_TOL).all()


def test_backward_rejects_equirect(device):
    r = Renderer(device)
    sc = Scene(device, S.gaussian_cloud(100, seed=1))
    try:
        with pytest.raises(Exception):
            r.forward(sc, S.equirect_camera(64, 32))
        with pytest.raises(Exception):
            r.backward(np.zeros((32, 64, 3), np.float32))
    finally:
        sc.destroy()
        r.destroy()


def test_backward_after_another_render_is_rejected(device):
    """A render after the forward replaces the lists, records and accumulator
    the backward reads, so the backward must refuse rather than differentiate
    a different frame; so must a backward whose forward scene was destroyed."""
    from droplet_amd._lib import DropletError
    r = Renderer(device)
    sc = Scene(device, S.gaussian_cloud(2000, seed=5, sh_bands=1))
    try:
        small, big = S.default_camera(96, 64), S.default_camera(256, 160)
        img = r.forward(sc, small)
        g = r.backward(np.ones_like(img))          # valid right after the forward
        assert np.isfinite(g["positions"]).all()
        r.forward(sc, small)
        r.render_scene(sc, big)
        with pytest.raises(DropletError, match="dl_render_forward"):
            r.backward(np.ones((big.height, big.width, 3), np.float32))
        img = r.forward(sc, small)
        r.autotune(sc, big)
        with pytest.raises(DropletError, match="dl_render_forward"):
            r.backward(np.ones_like(img))
        other = Scene(device, S.gaussian_cloud(500, seed=6, sh_bands=1))
        img = r.forward(other, small)
        other.destroy()
        with pytest.raises(DropletError, match="destroyed"):
            r.backward(np.ones_like(img))
    finally:
        sc.destroy()
        r.destroy()
