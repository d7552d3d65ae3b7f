"""Seeded synthetic scenes and cameras for the five benchmark workloads, plus
the deterministic fixtures of the reference's browser test page.

Every generator draws in float64 from ``numpy.random.default_rng(seed)`` and
stores float32, so the same seed gives the same bytes on every machine that
has this numpy. Readings of ambiguous distribution text:

* ``N(0, s^2)`` means standard deviation ``s``; ``N(0, 2)`` for opacity
  logits means standard deviation 2.
* "config-0 distributions scaled to radius 3" means positions ~ N(0, 0.75^2)
  per axis clipped to [-3, 3].
* The voxelizer box is [-5, 5] x [0, 3] x [-5, 5] metres.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .math import mat4_identity, mat4_look_at
from .types import RenderCamera, SplatData, num_sh_coeffs_per_channel


def _unit_quats(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32).reshape(-1)


def gaussian_cloud(n: int, seed: int = 0, *, pos_std: float = 0.5, pos_clip: float = 2.0,
                   log_scale_range: Tuple[float, float] = (math.log(0.003), math.log(0.05)),
                   sh_bands: int = 0, sh_std: float = 0.1) -> SplatData:
    """Render workload distributions: positions ~N(0, pos_std^2) per axis
    clipped to [-pos_clip, pos_clip]; rotations normalised N(0,1)^4 [w,x,y,z];
    logScales ~U(range); opacityLogits ~N(0, 2^2); colorsDC ~U[-1.5, 1.5];
    SH coefficients ~N(0, sh_std^2)."""
    rng = np.random.default_rng(seed)
    pos = np.clip(rng.normal(0.0, pos_std, (n, 3)), -pos_clip, pos_clip).astype(np.float32).reshape(-1)
    rot = _unit_quats(rng, n)
    ls = rng.uniform(log_scale_range[0], log_scale_range[1], (n, 3)).astype(np.float32).reshape(-1)
    op = rng.normal(0.0, 2.0, n).astype(np.float32)
    dc = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32).reshape(-1)
    m = num_sh_coeffs_per_channel(sh_bands)
    sh = rng.normal(0.0, sh_std, (n, 3 * m)).astype(np.float32).reshape(-1) if m else None
    return SplatData(n, pos, rot, ls, op, dc, sh, sh_bands)


def default_camera(width: int = 1280, height: int = 720) -> RenderCamera:
    """lookAt (0,0,4) -> origin, up +Y, fovY pi/3, near 0.01, far 1000."""
    return RenderCamera(width, height, mat4_look_at((0, 0, 4), (0, 0, 0), (0, 1, 0)),
                        math.pi / 3, 0.01, 1000.0, "pinhole")


def plain_render_workload(seed: int = 0):
    """100k Gaussians, bands 0, 1280x720 pinhole."""
    return gaussian_cloud(100_000, seed), default_camera(1280, 720)


def voxel_workload(n: int = 1_000_000, seed: int = 1):
    """Voxelizer workload: uniform positions in [-5,5]x[0,3]x[-5,5], random
    unit quaternions, logScales ~U[ln .005, ln .08], logits ~N(0, 2^2);
    voxelResolution 0.02, opacityCutoff 0.1."""
    rng = np.random.default_rng(seed)
    lo = np.array([-5.0, 0.0, -5.0])
    hi = np.array([5.0, 3.0, 5.0])
    pos = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32).reshape(-1)
    rot = _unit_quats(rng, n)
    ls = rng.uniform(math.log(0.005), math.log(0.08), (n, 3)).astype(np.float32).reshape(-1)
    op = rng.normal(0.0, 2.0, n).astype(np.float32)
    dc = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32).reshape(-1)
    return SplatData(n, pos, rot, ls, op, dc, None, 0), {"voxel_resolution": 0.02, "opacity_cutoff": 0.1}


def large_render_workload(n: int = 6_000_000, seed: int = 2):
    """6M Gaussians, positions ~N(0, 0.75^2) clipped to [-3, 3], shBands 3 with
    45 coefficients ~N(0, 0.1^2); 3840x2160 pinhole (default lookAt)."""
    return (gaussian_cloud(n, seed, pos_std=0.75, pos_clip=3.0, sh_bands=3),
            default_camera(3840, 2160))


def equirect_camera(width: int = 4096, height: int = 2048) -> RenderCamera:
    """Equirect from the origin (view = identity)."""
    return RenderCamera(width, height, mat4_identity(), math.pi / 3, 0.01, 1000.0, "equirect")


BATCH_SCENE_SIZES = (50_000, 200_000, 500_000, 1_000_000, 1_500_000, 2_000_000, 3_000_000, 4_000_000)
BATCH_RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080))


def orbit_cameras(count: int = 64, radius: float = 4.0) -> List[RenderCamera]:
    """Cameras on a radius-4 orbit in the y=0 plane looking at the origin;
    resolution cycles 640x360, 1280x720, 1920x1080 by camera index."""
    cams = []
    for k in range(count):
        a = 2.0 * math.pi * k / count
        eye = (radius * math.sin(a), 0.0, radius * math.cos(a))
        w, h = BATCH_RESOLUTIONS[k % 3]
        cams.append(RenderCamera(w, h, mat4_look_at(eye, (0, 0, 0), (0, 1, 0)), math.pi / 3, 0.01, 1000.0))
    return cams


def batch_workload(cameras: int = 64):
    """Batched-views workload: 8 scenes with the render distributions sized
    BATCH_SCENE_SIZES (seeds 100..107) and `cameras` orbit cameras shared by
    all scenes. Returns (scenes' SplatData, views), views scene-major: view
    si * cameras + k is scene si seen by camera k."""
    from .batch import View
    datas = [gaussian_cloud(n, seed=100 + i) for i, n in enumerate(BATCH_SCENE_SIZES)]
    cams = orbit_cameras(cameras)
    return datas, [View(si, cam) for si in range(len(datas)) for cam in cams]


# ---------------------------------------------------------------------------
# Browser test page fixtures (examples/browser/test.ts:45-103)
# ---------------------------------------------------------------------------
def one_splat(pos, color=(1.0, 0.0, 0.0), log_scale: float = -2.5, op_logit: float = 3.0) -> SplatData:
    """`oneSplat` (test.ts:45-56)."""
    return SplatData(1, np.array(pos, np.float32), np.array([1, 0, 0, 0], np.float32),
                     np.full(3, log_scale, np.float32), np.array([op_logit], np.float32),
                     np.array([color[0] - 0.5, color[1] - 0.5, color[2] - 0.5], np.float32))


def many_splats(n: int, layout: str = "sphere") -> SplatData:
    """`manySplats` (test.ts:58-93) with its u32 LCG
    seed = (seed*1664525 + 1013904223) | 0, value = (seed>>>0 % 10000)/10000."""
    positions = np.zeros(3 * n, np.float32)
    rotations = np.zeros(4 * n, np.float32)
    log_scales = np.full(3 * n, -3.2, np.float32)
    opacity = np.full(n, 2.5, np.float32)
    dc = np.zeros(3 * n, np.float32)
    seed = 1

    def rand():
        nonlocal seed
        seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
        return (seed % 10000) / 10000.0

    golden = math.pi * (1 + math.sqrt(5))
    for i in range(n):
        if layout == "sphere":
            t = i / n
            phi = math.acos(1 - 2 * t)
            theta = golden * i
            positions[3 * i:3 * i + 3] = (math.cos(theta) * math.sin(phi), math.sin(theta) * math.sin(phi), math.cos(phi))
        elif layout == "random":
            positions[3 * i] = (rand() - 0.5) * 4
            positions[3 * i + 1] = (rand() - 0.5) * 4
            positions[3 * i + 2] = (rand() - 0.5) * 4
        else:
            positions[3 * i + 2] = (i / n) * 2 - 1
        rotations[4 * i] = 1.0
        dc[3 * i] = rand() - 0.5
        dc[3 * i + 1] = rand() - 0.5
        dc[3 * i + 2] = rand() - 0.5
    return SplatData(n, positions, rotations, log_scales, opacity, dc, None, 0)


def dense_cluster(n: int = 200_000, seed: int = 7) -> SplatData:
    """Large-pair-count case (test.ts:215-243): a 0.1 m cube of splats with
    logScale -2.5 and logit 0. The page uses unseeded Math.random; a seeded
    generator stands in."""
    rng = np.random.default_rng(seed)
    pos = ((rng.random((n, 3)) - 0.5) * 0.1).astype(np.float32).reshape(-1)
    rot = np.zeros((n, 4), np.float32); rot[:, 0] = 1
    ls = np.full(3 * n, -2.5, np.float32)
    op = np.zeros(n, np.float32)
    dc = np.zeros((n, 3), np.float32); dc[:, 0] = 0.5
    return SplatData(n, pos, rot.reshape(-1), ls, op, dc.reshape(-1), None, 0)


def empty_scene() -> SplatData:
    z = np.zeros(0, np.float32)
    return SplatData(0, z, z, z, z, z, None, 0)
