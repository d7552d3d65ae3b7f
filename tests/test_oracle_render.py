"""Pins the oracle's render pipeline on the CPU.

What each test fixes (a plausible slip in the oracle fails at least one):
* single-splat closed forms: EWA covariance + 0.3 dilation, Mip-Splatting
  alpha factor, conic, 3-sigma radius, sigmoid opacity, colour offset, and
  the composited pixel value as a function of distance;
* the sign-flipped cov2_01 of project.ts:128-133 against a textbook EWA
  evaluation in f64;
* the SH basis (bands 1-3) against the real spherical-harmonic polynomials
  evaluated in f64, one basis function at a time;
* an untiled brute-force rasterizer (every pixel walks all depth-sorted
  splats) equals the tiled pipeline bit for bit, which pins emit, the stable
  pair sort, tile ranges and the T early-out;
* pair lists against an independent enumeration of each splat's tile
  rectangle (including the row 0 / column 0 clamp quirk and maxCoverage
  truncation), and tile ranges against searchsorted;
* chunked rendering is bit-identical to one chunk;
* the sampled-row mode (only some tile rows emitted and rasterized) gives
  those rows bit-identical to the full render, the full pair count, and
  per-row loop counts that add up to the full render's;
* the equirect screen mapping against its closed form in every direction
  (ahead, +-x, up, down, both sides of the seam), and the equirect path's
  reuse of the pinhole Jacobian (records equal to the pinhole projection's
  apart from the screen position);
* the reference's browser test page cases (examples/browser/test.ts:114-251).
"""
import math

import numpy as np
import pytest

import oracle
from droplet_amd import synthetic as S
from droplet_amd.math import mat4_look_at
from droplet_amd.types import BackgroundRGBA, RenderCamera, SplatData

SH1 = 0.4886025119029199


def _sigmoid(x):
    return 1.0 / (1.0 + math.exp(-x))


def _single_splat_expect(ls, logit, W, H, fov=math.pi / 3):
    f = H / (2 * math.tan(fov / 2))
    s = math.exp(ls)
    sp2 = (f * s / 4.0) ** 2
    sb2 = sp2 + 0.3
    aa = sp2 / sb2
    radius = math.ceil(3 * math.sqrt(sb2 + math.sqrt(0.1)))
    alpha = _sigmoid(logit) * aa
    return f, sb2, radius, alpha


@pytest.mark.parametrize("ls,logit", [(-2.5, 3.0), (-4.0, 0.5), (-1.5, -1.0), (-6.0, 2.0)])
def test_single_splat_projection_closed_form(ls, logit):
    W, H = 256, 192
    data = S.one_splat((0, 0, 0), (0.9, 0.3, 0.1), ls, logit)
    cam = S.default_camera(W, H)
    r = oracle.render(data, cam)
    rec = r["projected"][0].astype(np.float64)
    f, sb2, radius, alpha = _single_splat_expect(ls, logit, W, H)
    assert rec[0] == W / 2 and rec[1] == H / 2
    assert rec[2] == radius
    assert rec[3] == 4.0
    assert rec[4] == pytest.approx(1 / sb2, rel=2e-6)
    assert rec[5] == 0.0
    assert rec[6] == pytest.approx(1 / sb2, rel=2e-6)
    assert rec[7] == pytest.approx(alpha, rel=2e-6)
    np.testing.assert_allclose(rec[8:11], [0.9, 0.3, 0.1], rtol=1e-6)


def test_single_splat_image_closed_form():
    W, H = 128, 128
    ls, logit = -2.5, 3.0
    bg = BackgroundRGBA(0.1, 0.2, 0.3, 1.0)
    data = S.one_splat((0, 0, 0), (1.0, 0.0, 0.0), ls, logit)
    r = oracle.render(data, S.default_camera(W, H), background=bg)
    img = r["rgba"].reshape(H, W, 4).astype(np.int64)
    f, sb2, radius, alpha = _single_splat_expect(ls, logit, W, H)
    floor = float(np.float32(1.1108996538e-2))
    ys, xs = np.mgrid[0:H, 0:W]
    dx = xs + 0.5 - W / 2
    dy = ys + 0.5 - H / 2
    a = np.minimum(0.999, alpha * np.maximum(0.0, np.exp(-0.5 * (dx * dx + dy * dy) / sb2) - floor))
    a = np.where((np.abs(dx) > radius) | (np.abs(dy) > radius) | (a < 1 / 255), 0.0, a)
    T = 1 - a
    exp_r = np.clip(a * 1.0 + T * 0.1, 0, 1) * 255 + 0.5
    exp_g = np.clip(T * 0.2, 0, 1) * 255 + 0.5
    # 1 LSB slack only where the f64 value sits on a rounding boundary
    assert np.abs(img[..., 0] - np.floor(exp_r)).max() <= 1
    assert np.abs(img[..., 1] - np.floor(exp_g)).max() <= 1
    assert (np.abs(img[..., 0] - np.floor(exp_r)) > 0).mean() < 0.01
    # exact mirror symmetry about the image centre (sx = W/2 exactly)
    np.testing.assert_array_equal(img, img[:, ::-1])
    np.testing.assert_array_equal(img, img[::-1, :])
    # reference test page: centre red > 100, |L-R| < 10 at +-10 px (test.ts:126-140)
    assert img[64, 64, 0] > 100


def test_cov01_sign_flip_against_textbook_ewa():
    """project.ts:128-133 negates row 0 of the EWA Jacobian, so the conic's
    B term has the opposite sign of the textbook (screen-space) result."""
    rng = np.random.default_rng(0)
    W, H = 800, 600
    cam = S.default_camera(W, H)
    f = oracle.focal(cam)
    for _ in range(20):
        pos = rng.uniform(-0.8, 0.8, 3)
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        ls = rng.uniform(-4, -2, 3)
        rec = np.concatenate([pos, q, ls, [1.0], [0.0, 0.0, 0.0]]).astype(np.float32)
        out, _ = oracle.project_packed(rec, 0, cam)
        w, x, y, z = q.astype(np.float32).astype(np.float64)
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        Sg = np.diag(np.exp(ls.astype(np.float32).astype(np.float64)) ** 2)
        V = cam.view_matrix.astype(np.float64).reshape(4, 4).T
        pc = V[:3, :3] @ pos.astype(np.float32).astype(np.float64) + V[:3, 3]
        Sc = V[:3, :3] @ R @ Sg @ R.T @ V[:3, :3].T
        zc = -pc[2]
        J = np.array([[f / zc, 0, f * pc[0] / zc ** 2],       # d(screenX)/d(posCam), screenX = x f / -z
                      [0, -f / zc, -f * pc[1] / zc ** 2]])    # screenY = -y f / -z
        cov = J @ Sc @ J.T
        cov[0, 0] += 0.3
        cov[1, 1] += 0.3
        conic = np.linalg.inv(cov)
        assert out[4] == pytest.approx(conic[0, 0], rel=1e-4)
        assert out[6] == pytest.approx(conic[1, 1], rel=1e-4)
        # textbook off-diagonal is conic[0,1]; the reference stores its negation
        assert out[5] == pytest.approx(-conic[0, 1], rel=1e-3, abs=1e-7)
        assert out[0] == pytest.approx(pc[0] * f / zc + W / 2, rel=1e-5)
        assert out[1] == pytest.approx(-pc[1] * f / zc + H / 2, rel=1e-5)


def test_equirect_screen_mapping_closed_form():
    """Equirect screen position (project.ts:251-262): a splat in camera-space
    direction (lon, lat), with +x right, +y up and -z forward, lands at
    sx = (lon / (2 P) + 0.5) W and sy = (0.5 - lat / P) H, P = 3.14159265
    (the shader's literal, not pi). Covers straight ahead, towards +-x, up
    and down, both sides of the seam behind the camera and generic
    directions; a mirrored or swapped axis moves the splat by a large
    fraction of the image. Directions stay away from camera-space z = 0,
    where the reused pinhole Jacobian (project.ts:128-133) makes the
    covariance blow up and the fade culls the splat."""
    W, H = 4096, 2048
    cam = S.equirect_camera(W, H)
    P = 3.14159265
    r = 3.0
    dirs = [(0.0, 0.0), (1.45, 0.0), (-1.45, 0.0), (0.0, 1.4), (0.0, -1.4), (math.pi - 0.05, 0.1),
            (-math.pi + 0.05, -0.1), (0.7, 0.4), (-2.1, -0.6), (2.9, 1.2), (-0.3, -1.3)]
    for lon, lat in dirs:
        pos = r * np.array([math.sin(lon) * math.cos(lat), math.sin(lat), -math.cos(lon) * math.cos(lat)])
        rec = np.concatenate([pos, [1, 0, 0, 0], [-9, -9, -9], [2.0], [0.1, 0.2, 0.3]]).astype(np.float32)
        out, _ = oracle.project_packed(rec, 0, cam)
        assert out[2] > 0, (lon, lat)  # not culled
        p = rec[:3].astype(np.float64)
        assert abs(math.atan2(p[0], -p[2]) - lon) < 1e-6 and abs(math.asin(p[1] / np.linalg.norm(p)) - lat) < 1e-6
        assert out[0] == pytest.approx((lon / (2 * P) + 0.5) * W, abs=2e-2), (lon, lat)
        assert out[1] == pytest.approx((0.5 - lat / P) * H, abs=2e-2), (lon, lat)

    def at(x, y, z):
        rec = np.array([x, y, z, 1, 0, 0, 0, -9, -9, -9, 2.0, 0, 0, 0], np.float32)
        out = oracle.project_packed(rec, 0, cam)[0]
        assert out[2] > 0
        return out[:2]
    # the named directions as plain facts about the image
    assert at(0, 0, -2) == pytest.approx([W / 2, H / 2], abs=1e-3)     # ahead: the centre
    assert 0.7 * W < at(2, 0, -0.2)[0] < 0.75 * W                    # right of ahead: towards 3/4
    assert 0.25 * W < at(-2, 0, -0.2)[0] < 0.3 * W                   # left: towards 1/4
    assert at(0, 2, -0.2)[1] < 0.05 * H and at(0, -2, -0.2)[1] > 0.95 * H  # up: top rows; down: bottom
    assert at(0.01, 0, 2)[0] > 0.99 * W and at(-0.01, 0, 2)[0] < 0.01 * W  # behind: the seam at the edges
    # straight behind with x = +0: atan2(+0, -z) = +pi, just past the right edge
    assert at(0, 0, 2)[0] == pytest.approx((math.pi / (2 * P) + 0.5) * W, abs=1e-2)


def test_equirect_reuses_the_pinhole_jacobian():
    """project.ts:125-149 runs after either projection body, so an equirect
    splat's conic, radius, alpha and colour are those of the pinhole
    projection of the same camera-space point (same focal, same height):
    only the screen position differs."""
    rng = np.random.default_rng(11)
    W, H = 1024, 512
    eq = S.equirect_camera(W, H)
    ph = RenderCamera(W, H, eq.view_matrix.copy(), eq.fov_y, eq.near, eq.far, "pinhole")
    checked = 0
    for _ in range(40):
        pos = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -rng.uniform(1.5, 4)])
        q = rng.standard_normal(4)
        rec = np.concatenate([pos, q / np.linalg.norm(q), rng.uniform(-4, -2, 3), [1.0], [0.1, 0.2, 0.3]])
        rec = rec.astype(np.float32)
        a, _ = oracle.project_packed(rec, 0, eq)
        b, _ = oracle.project_packed(rec, 0, ph)
        if b[2] <= 0:
            continue
        np.testing.assert_array_equal(a[2:], b[2:])
        assert not np.array_equal(a[:2], b[:2])
        checked += 1
    assert checked >= 30


def _sh_basis(d):
    """Real SH basis for l = 1..3 in the 3DGS sign convention, f64."""
    x, y, z = d
    return np.array([
        -SH1 * y, SH1 * z, -SH1 * x,
        1.0925484305920792 * x * y, -1.0925484305920792 * y * z,
        0.31539156525252005 * (2 * z * z - x * x - y * y), -1.0925484305920792 * x * z,
        0.5462742152960396 * (x * x - y * y),
        -0.5900435899266435 * y * (3 * x * x - y * y), 2.890611442640554 * x * y * z,
        -0.4570457994644658 * y * (4 * z * z - x * x - y * y),
        0.3731763325901154 * z * (2 * z * z - 3 * x * x - 3 * y * y),
        -0.4570457994644658 * x * (4 * z * z - x * x - y * y),
        1.445305721320277 * z * (x * x - y * y), -0.5900435899266435 * x * (x * x - 3 * y * y)])


def test_sh_basis_functions_one_at_a_time():
    """Each coefficient k of channel c contributes basis_k(viewDir) to channel
    c only, with viewDir = normalize(posCam) (camera space, project.ts:206)."""
    rng = np.random.default_rng(1)
    cam = S.default_camera(512, 512)
    for bands, M in [(1, 3), (2, 8), (3, 15)]:
        for k in range(M):
            for c in range(3):
                pos = rng.uniform(-1, 1, 3)
                sh = np.zeros(3 * M)
                sh[c * M + k] = 0.25
                rec = np.concatenate([pos, [1, 0, 0, 0], [-3, -3, -3], [2.0], [0.1, 0.2, 0.3], sh]).astype(np.float32)
                out, _ = oracle.project_packed(rec, bands, cam)
                p32 = rec[:3].astype(np.float64)
                pc = p32 + np.array([0, 0, -4.0])
                d = pc / np.linalg.norm(pc)
                want = np.array([0.1, 0.2, 0.3]) + 0.5
                want[c] += 0.25 * _sh_basis(d)[k]
                np.testing.assert_allclose(out[8:11], np.maximum(want, 0), rtol=0, atol=2e-6)


def _brute_force_accum(proj, W, H):
    """Untiled rasterizer: every pixel walks every visible splat in depth
    order with the reference's per-pixel body (rasterize.ts:62-79), f32."""
    f32 = np.float32
    ys, xs = np.mgrid[0:H, 0:W]
    px = xs.astype(f32) + f32(0.5)
    py = ys.astype(f32) + f32(0.5)
    cr = np.zeros((H, W), f32); cg = np.zeros((H, W), f32); cb = np.zeros((H, W), f32)
    T = np.ones((H, W), f32)
    cap, tmin, amin = f32(0.999), f32(0.0001), f32(0.00392156862745098)
    floor = f32(1.1108996538e-2)
    for rec in proj:
        sx, sy, r = rec[0], rec[1], rec[2]
        if not r > 0:
            continue
        A, B, C, a = rec[4], rec[5], rec[6], rec[7]
        live = ~(T < tmin)
        dx = px - sx
        dy = py - sy
        inbox = live & ~((np.abs(dx) > r) | (np.abs(dy) > r))
        # rasterize.ts:71 with the specification's two contractions
        power = f32(-0.5) * oracle.fmaf(C * dy, dy, oracle.fmaf((f32(2.0) * B) * dx, dy, (A * dx) * dx))
        e = oracle.expf(power).reshape(H, W)
        alpha = np.minimum(cap, a * np.maximum(f32(0.0), e - floor))
        upd = inbox & ~(power > 0) & ~(alpha < amin)
        w = T * alpha
        cr = np.where(upd, oracle.fmaf(w, rec[8], cr), cr)
        cg = np.where(upd, oracle.fmaf(w, rec[9], cg), cg)
        cb = np.where(upd, oracle.fmaf(w, rec[10], cb), cb)
        T = np.where(upd, T * (f32(1.0) - alpha), T)
    return np.stack([cr, cg, cb, T], -1)


@pytest.mark.parametrize("seed,n,W,H,bands", [(0, 400, 96, 64, 0), (1, 300, 80, 48, 2), (2, 500, 70, 70, 1)])
def test_tiled_equals_untiled_brute_force(seed, n, W, H, bands):
    data = S.gaussian_cloud(n, seed=seed, sh_bands=bands, log_scale_range=(math.log(0.01), math.log(0.15)))
    cam = S.default_camera(W, H)
    r = oracle.render(data, cam)
    assert r["num_pairs"] > n  # multi-tile splats present
    brute = _brute_force_accum(r["projected"], W, H)
    np.testing.assert_array_equal(r["accum"].view(np.uint32), brute.view(np.uint32))


def _expected_pairs(proj, W, H, max_cov):
    tx_n, ty_n = (W + 15) // 16, (H + 15) // 16
    f32 = np.float32
    tiles, splats = [], []
    for j, rec in enumerate(proj):
        sx, sy, r = f32(rec[0]), f32(rec[1]), f32(rec[2])
        if not r > 0:
            continue
        def sat(v):
            return 0 if not v > 0 else int(v)
        x0 = sat(max(f32(0), np.floor((sx - r) / f32(16))))
        y0 = sat(max(f32(0), np.floor((sy - r) / f32(16))))
        x1 = sat(min(f32(tx_n) - f32(1), np.floor((sx + r) / f32(16))))
        y1 = sat(min(f32(ty_n) - f32(1), np.floor((sy + r) / f32(16))))
        written = 0
        for ty in range(y0, y1 + 1):
            for tx in range(x0, x1 + 1):
                if written >= max_cov:
                    break
                tiles.append(ty * tx_n + tx)
                splats.append(j)
                written += 1
    tiles = np.array(tiles, np.uint32)
    splats = np.array(splats, np.uint32)
    order = np.argsort(tiles, kind="stable")
    return tiles[order], splats[order]


@pytest.mark.parametrize("max_cov", [0, 3])
def test_pairs_and_ranges_against_rectangle_enumeration(max_cov):
    # splats off the left/top edge still land in column/row 0 (u32 clamp quirk)
    data = S.gaussian_cloud(3000, seed=4, pos_std=1.2, pos_clip=3.0)
    W, H = 200, 120
    cam = S.default_camera(W, H)
    r = oracle.render(data, cam, max_coverage=max_cov)
    et, es = _expected_pairs(r["projected"], W, H, max_cov or 16384)
    np.testing.assert_array_equal(r["pair_tiles"], et)
    np.testing.assert_array_equal(r["pair_splats"], es)
    T = r["num_tiles"]
    off = r["tile_offsets"][0]
    np.testing.assert_array_equal(off, np.searchsorted(et, np.arange(T + 1), side="left"))
    counts = np.bincount(es, minlength=r["num_visible"])
    np.testing.assert_array_equal(counts, r["tile_counts"])
    # the clamp quirk is exercised: some splat left of the image sits in column 0
    proj = r["projected"]
    left = (proj[:, 2] > 0) & (proj[:, 0] + proj[:, 2] < 0)
    assert left.any()


def test_chunked_render_is_bit_identical():
    data = S.gaussian_cloud(40_000, seed=5)
    cam = S.default_camera(160, 90)
    one = oracle.render(data, cam, chunk_cap=10 ** 9)
    many = oracle.render(data, cam, chunk_cap=9_000)
    assert many["num_chunks"] == 5
    np.testing.assert_array_equal(one["rgba"], many["rgba"])
    np.testing.assert_array_equal(one["accum"].view(np.uint32), many["accum"].view(np.uint32))
    assert one["num_pairs"] == many["num_pairs"]


# --- the reference's browser test page (examples/browser/test.ts:114-251) ---
def _nonblack(rgba):
    p = rgba.reshape(-1, 4)
    return int(((p[:, 0] > 5) | (p[:, 1] > 5) | (p[:, 2] > 5)).sum())


def test_browser_empty_scene_all_black():
    r = oracle.render(S.empty_scene(), S.default_camera(128, 128), taps=False)
    assert _nonblack(r["rgba"]) == 0


def test_browser_three_splats():
    data = SplatData(3, [0, 0, 1, 0.5, 0, 0.8, -0.5, 0.3, 0.8], [1, 0, 0, 0] * 3, [-2.5] * 9, [3, 3, 3],
                     [0.5, -0.5, -0.5, -0.5, 0.5, -0.5, -0.5, -0.5, 0.5])
    img = oracle.render(data, S.default_camera(256, 256), taps=False)["rgba"].reshape(-1, 4).astype(int)
    r, g, b = img[:, 0], img[:, 1], img[:, 2]
    assert ((r > 50) & (r > g) & (r > b)).sum() > 50
    assert ((g > 50) & (g > r) & (g > b)).sum() > 50
    assert ((b > 50) & (b > r) & (b > g)).sum() > 50


def test_browser_depth_occlusion():
    data = SplatData(2, [0, 0, 1, 0, 0, -1], [1, 0, 0, 0] * 2, [-2.5] * 6, [4, 4], [0.5, -0.5, -0.5, -0.5, 0.5, -0.5])
    img = oracle.render(data, S.default_camera(128, 128), taps=False)["rgba"].reshape(128, 128, 4)
    assert img[64, 64, 0] > 200 and img[64, 64, 1] < 50


def test_browser_sphere_disc():
    img = oracle.render(S.many_splats(5000, "sphere"), S.default_camera(256, 256), taps=False)["rgba"]
    assert 10_000 < _nonblack(img) < 30_000


def test_browser_large_pair_count():
    img = oracle.render(S.dense_cluster(200_000), S.default_camera(256, 256), taps=False)["rgba"]
    assert _nonblack(img) > 100


def test_browser_chunk_boundary():
    data = S.many_splats(200_000, "random")
    cam = S.default_camera(64, 64)
    a = oracle.render(data, cam, taps=False)["rgba"]
    b = oracle.render(data, cam, chunk_cap=50_000, taps=False)["rgba"]
    assert np.abs(a.astype(int) - b.astype(int)).max() <= 2


@pytest.mark.parametrize("step,phase", [(1, 0), (3, 1), (5, 4), (16, 0)])
def test_sampled_rows_equal_the_full_render(step, phase):
    """orc_render_rows keeps only the pairs of every step-th tile row: the
    stable pair sort keeps each tile's list order, and maxCoverage counts
    the dropped pairs too, so the kept rows match the full render bit for
    bit. Covers maxCoverage truncation and several reference chunks."""
    data = S.gaussian_cloud(30_000, seed=9, pos_std=0.6, log_scale_range=(math.log(0.005), math.log(0.2)))
    cam = S.default_camera(400, 300)
    bg = BackgroundRGBA(0.2, 0.1, 0.3, 0.8)
    for chunk, cov in ((0, 0), (7_000, 23)):
        full = oracle.render(data, cam, background=bg, chunk_cap=chunk, max_coverage=cov, taps=True)
        part = oracle.render_rows(data, cam, background=bg, row_step=step, row_phase=phase, chunk_cap=chunk,
                                  max_coverage=cov)
        assert part["num_pairs"] == full["num_pairs"] and part["num_visible"] == full["num_visible"]
        a = full["rgba"].reshape(cam.height, cam.width, 4)
        b = part["rgba"].reshape(cam.height, cam.width, 4)
        ty = np.arange(cam.height) // 16
        kept = (ty % step) == (phase % step)
        np.testing.assert_array_equal(b[kept], a[kept])
        # rows not kept show the background only
        f = np.float32
        bgv = np.array([0.2, 0.1, 0.3, (f(1) - f(1)) + f(1) * f(0.8)], np.float32)
        bgpx = (bgv * f(255) + f(0.5)).astype(np.uint32).astype(np.uint8)  # finalize.ts:35-40
        assert (b[~kept] == bgpx).all()
        if step == 1:
            assert int(part["row_iters"].sum()) == full["pixel_iters"]
        else:
            one = oracle.render_rows(data, cam, background=bg, row_step=1, chunk_cap=chunk, max_coverage=cov)
            np.testing.assert_array_equal(part["row_iters"][part["tile_rows"]], one["row_iters"][part["tile_rows"]])
            assert (np.delete(part["row_iters"], part["tile_rows"]) == 0).all()


def test_chunk_over_u32_pairs_is_an_error_not_a_wrap():
    """A chunk whose pair total needs more than 32 bits is refused, not
    wrapped (the reference's pair buffers are indexed by u32): 270k splats
    that each cover the whole 2048x2048 frame are 270k x 16384 > 2^32 pairs
    in one chunk."""
    n = 270_000
    rng = np.random.default_rng(3)
    pos = np.zeros((n, 3), np.float32)
    pos[:, :2] = rng.uniform(-0.05, 0.05, (n, 2))
    pos[:, 2] = 3.0
    data = SplatData(n, pos.reshape(-1), np.tile([1, 0, 0, 0], n).astype(np.float32),
                     np.full(3 * n, -0.5, np.float32), np.full(n, -6.0, np.float32), np.zeros(3 * n, np.float32))
    cam = S.default_camera(2048, 2048)
    with pytest.raises(RuntimeError, match=r"\(-2\)"):
        oracle.render_rows(data, cam, row_step=64, chunk_cap=0xFFFFFFFF)
