"""Host-side camera math (src/math/index.ts).

Matrices are column-major float32[16], computed in f64 and stored as f32 the
way the reference writes a Float32Array.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np


def vec3(x: float, y: float, z: float):
    return (float(x), float(y), float(z))


def _norm(v):
    length = math.hypot(v[0], v[1], v[2])
    return (v[0] / length, v[1] / length, v[2] / length) if length > 0 else (0.0, 0.0, 0.0)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def mat4_identity() -> np.ndarray:
    m = np.zeros(16, dtype=np.float32)
    m[0] = m[5] = m[10] = m[15] = 1.0
    return m


def mat4_look_at(eye: Sequence[float], target: Sequence[float], up: Sequence[float]) -> np.ndarray:
    """Right-handed look-at, world->camera, camera looking down -z
    (`mat4LookAt`, src/math/index.ts:41-56)."""
    f = _norm((target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]))
    s = _norm(_cross(f, up))
    u = _cross(s, f)
    m = np.zeros(16, dtype=np.float64)
    m[0], m[1], m[2], m[3] = s[0], u[0], -f[0], 0.0
    m[4], m[5], m[6], m[7] = s[1], u[1], -f[1], 0.0
    m[8], m[9], m[10], m[11] = s[2], u[2], -f[2], 0.0
    m[12] = -_dot(s, eye)
    m[13] = -_dot(u, eye)
    m[14] = _dot(f, eye)
    m[15] = 1.0
    return m.astype(np.float32)
