"""Data types of droplet's public API, mirrored for the HIP build.

`SplatData` follows `src/types.ts:23-32`, `RenderCamera` `src/types.ts:74-82`,
`BackgroundRGBA` `src/types.ts:60-65`, `Bounds` `src/types.ts:54-57`.
Arrays are float32 numpy arrays in the reference's layouts:

* positions 3N, rotations 4N ``[w, x, y, z]``, log_scales 3N,
  opacity_logits N, colors_dc 3N,
* colors_sh 3*M*N, per splat channel-major ``R[0..M) G[0..M) B[0..M)``
  with M = num_sh_coeffs_per_channel(sh_bands) (`src/types.ts:35-37`).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

Vec3 = Tuple[float, float, float]


def num_sh_coeffs_per_channel(bands: int) -> int:
    """`numSHCoeffsPerChannel` (src/types.ts:35-37)."""
    if bands not in (0, 1, 2, 3):
        raise ValueError(f"shBands must be 0..3, got {bands}")
    return (0, 3, 8, 15)[bands]


def splat_input_stride(bands: int) -> int:
    """`splatInputStride` (src/types.ts:49-51): floats per packed splat."""
    return 14 + 3 * num_sh_coeffs_per_channel(bands)


def _f32(a, n: int, name: str) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
    if arr.size != n:
        raise ValueError(f"{name}: expected {n} float32 values, got {arr.size}")
    return arr


@dataclass
class SplatData:
    """Struct-of-arrays Gaussian set (src/types.ts:23-32)."""

    count: int
    positions: np.ndarray
    rotations: np.ndarray
    log_scales: np.ndarray
    opacity_logits: np.ndarray
    colors_dc: np.ndarray
    colors_sh: Optional[np.ndarray] = None
    sh_bands: int = 0

    def __post_init__(self) -> None:
        n = int(self.count)
        if n < 0:
            raise ValueError("count must be >= 0")
        self.count = n
        self.positions = _f32(self.positions, 3 * n, "positions")
        self.rotations = _f32(self.rotations, 4 * n, "rotations")
        self.log_scales = _f32(self.log_scales, 3 * n, "logScales")
        self.opacity_logits = _f32(self.opacity_logits, n, "opacityLogits")
        self.colors_dc = _f32(self.colors_dc, 3 * n, "colorsDC")
        m = num_sh_coeffs_per_channel(self.sh_bands)
        if m > 0:
            if self.colors_sh is None:
                raise ValueError("colorsSH is required when shBands > 0")
            sh = np.ascontiguousarray(np.asarray(self.colors_sh, dtype=np.float32).reshape(-1))
            # loadPlySplats (src/io/ply.ts:181-188) can hold more coefficients
            # per splat than shBands uses; the renderer reads the first 3*M*N
            # values with stride 3*M, as the reference's packer does
            if sh.size < 3 * m * n:
                raise ValueError(f"colorsSH: expected at least {3 * m * n} float32 values, got {sh.size}")
            self.colors_sh = sh
        elif self.colors_sh is not None:
            self.colors_sh = np.ascontiguousarray(np.asarray(self.colors_sh, dtype=np.float32).reshape(-1))

    def nbytes(self) -> int:
        """Bytes of the arrays the renderer uploads (colorsSH up to 3*M*N)."""
        m = num_sh_coeffs_per_channel(self.sh_bands)
        return 4 * (14 * self.count + 3 * m * self.count)


@dataclass
class RenderCamera:
    """View parameters (src/types.ts:74-82). `view_matrix` is the column-major
    world->camera Float32Array(16); fov_y/near/far stay Python floats (JS
    numbers, f64) so the f64 cull of src/render/camera.ts:56-99 is kept."""

    width: int
    height: int
    view_matrix: np.ndarray
    fov_y: float
    near: float
    far: float
    projection: str = "pinhole"

    def __post_init__(self) -> None:
        self.width = int(self.width)
        self.height = int(self.height)
        self.view_matrix = _f32(self.view_matrix, 16, "viewMatrix")
        if self.projection not in ("pinhole", "equirect"):
            raise ValueError(f"projection must be 'pinhole' or 'equirect', got {self.projection!r}")


@dataclass
class BackgroundRGBA:
    """Background colour in [0, 1] (src/types.ts:60-65)."""

    r: float = 0.0
    g: float = 0.0
    b: float = 0.0
    a: float = 1.0


@dataclass
class Bounds:
    """Axis-aligned box (src/types.ts:54-57)."""

    min: Vec3 = field(default_factory=lambda: (0.0, 0.0, 0.0))
    max: Vec3 = field(default_factory=lambda: (0.0, 0.0, 0.0))
