"""`voxelize` and `BlockMaskBuffer`: the Python mirror of droplet's voxelizer
API (src/voxel/voxelize.ts:29-63, src/voxel/block-mask-buffer.ts:23-58). The
work runs in libdroplet_hip.so on the GPU."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib as L
from .types import Bounds, SplatData


@dataclass
class VoxelizeOptions:
    """`VoxelizeOptions` (voxelize.ts:29-36)."""

    voxel_resolution: float = 0.05
    opacity_cutoff: float = 0.1
    bounds: Optional[Bounds] = None


@dataclass
class VoxelBlock:
    """`VoxelBlock` (block-mask-buffer.ts:15-22)."""

    block_idx: int
    mask_lo: int
    mask_hi: int


def _popcount32(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint32)
    v = v - ((v >> 1) & 0x55555555)
    v = (v & 0x33333333) + ((v >> 2) & 0x33333333)
    return (((v + (v >> 4)) & 0x0F0F0F0F) * np.uint32(0x01010101)) >> 24


class BlockMaskBuffer:
    """Sparse blockIdx -> 64-bit mask map with insertion order
    (block-mask-buffer.ts:23-58). Stored as three parallel u32 arrays."""

    def __init__(self, block_idx=None, mask_lo=None, mask_hi=None):
        self.block_idx = np.zeros(0, np.uint32) if block_idx is None else np.asarray(block_idx, np.uint32)
        self.mask_lo = np.zeros(0, np.uint32) if mask_lo is None else np.asarray(mask_lo, np.uint32)
        self.mask_hi = np.zeros(0, np.uint32) if mask_hi is None else np.asarray(mask_hi, np.uint32)
        self._index: Optional[Dict[int, int]] = None

    def _ensure_index(self):
        if self._index is None:
            self._index = {int(b): i for i, b in enumerate(self.block_idx)}
        return self._index

    def add_block(self, block_idx: int, mask_lo: int, mask_hi: int) -> None:
        """`addBlock` (block-mask-buffer.ts:30-41): insert or OR-merge."""
        if mask_lo == 0 and mask_hi == 0:
            return
        idx = self._ensure_index()
        if block_idx in idx:
            i = idx[block_idx]
            self.mask_lo[i] |= np.uint32(mask_lo)
            self.mask_hi[i] |= np.uint32(mask_hi)
        else:
            idx[block_idx] = len(self.block_idx)
            self.block_idx = np.append(self.block_idx, np.uint32(block_idx))
            self.mask_lo = np.append(self.mask_lo, np.uint32(mask_lo))
            self.mask_hi = np.append(self.mask_hi, np.uint32(mask_hi))

    def get(self, block_idx: int) -> Optional[VoxelBlock]:
        i = self._ensure_index().get(int(block_idx))
        if i is None:
            return None
        return VoxelBlock(int(self.block_idx[i]), int(self.mask_lo[i]), int(self.mask_hi[i]))

    @property
    def blocks(self) -> List[VoxelBlock]:
        return [VoxelBlock(int(b), int(lo), int(hi)) for b, lo, hi in zip(self.block_idx, self.mask_lo, self.mask_hi)]

    def popcount(self) -> int:
        """Total set voxel bits (block-mask-buffer.ts:49-55)."""
        return int(_popcount32(self.mask_lo).sum() + _popcount32(self.mask_hi).sum())

    @property
    def count(self) -> int:
        return int(self.block_idx.size)


@dataclass
class VoxelizeResult:
    """`VoxelizeResult` (voxelize.ts:38-45). num_blocks_* are floats because
    the reference's are JS numbers (-Infinity for an empty scene)."""

    buffer: BlockMaskBuffer
    bounds: Bounds
    voxel_resolution: float
    num_blocks_x: float
    num_blocks_y: float
    num_blocks_z: float
    device_ms: float = 0.0


def _options_struct(opts: VoxelizeOptions) -> "L.DlVoxelOpts":
    o = L.DlVoxelOpts()
    o.voxel_resolution = float(opts.voxel_resolution)
    o.opacity_cutoff = float(opts.opacity_cutoff)
    if opts.bounds is not None:
        o.has_bounds = 1
        for k in range(3):
            o.bounds_min[k] = float(opts.bounds.min[k])
            o.bounds_max[k] = float(opts.bounds.max[k])
    return o


def _points(points) -> np.ndarray:
    pts = np.ascontiguousarray(points, np.float64)
    if pts.size % 3:
        raise ValueError("points must be (x, y, z) triples")
    return pts.reshape(-1, 3)


@dataclass
class VoxelMesh:
    """Quad mesh of a voxel grid's solid boundary: 4 vertices and 2 triangles
    per quad (dl_voxel_mesh_build)."""

    positions: np.ndarray   # float32 [4Q, 3]
    indices: np.ndarray     # uint32 [2Q, 3]
    device_ms: float = 0.0

    @property
    def quads(self) -> int:
        return self.positions.shape[0] // 4


class VoxelGrid:
    """A BlockMaskBuffer resident in HBM: the dense grid of 64-bit block masks
    that `voxelize` produces on the device, and the collision stages that read
    it there (fill, carve, mesh, query). The reference keeps the buffer on the
    host (src/voxel/block-mask-buffer.ts:23-58) and leaves these stages to
    callers (README.md:108); their definitions are in include/droplet_hip.h."""

    def __init__(self, device, handle, voxel_resolution: float):
        self._device = device
        self._h = handle
        self.voxel_resolution = float(voxel_resolution)

    @classmethod
    def from_result(cls, device, result: VoxelizeResult) -> "VoxelGrid":
        """Uploads a host VoxelizeResult (repeated blocks OR-merge as addBlock)."""
        buf = result.buffer
        lib = L.load()
        h = C.c_void_p()
        bmin = (C.c_double * 3)(*[float(v) for v in result.bounds.min])
        bmax = (C.c_double * 3)(*[float(v) for v in result.bounds.max])
        nb = (C.c_double * 3)(result.num_blocks_x, result.num_blocks_y, result.num_blocks_z)
        idx = np.ascontiguousarray(buf.block_idx, np.uint32)
        lo = np.ascontiguousarray(buf.mask_lo, np.uint32)
        hi = np.ascontiguousarray(buf.mask_hi, np.uint32)
        if not (idx.size == lo.size == hi.size):
            raise ValueError("block arrays differ in length")
        L.check(lib.dl_voxel_result_create(device.handle, bmin, bmax, nb, float(result.voxel_resolution),
                                           L.ptr(idx, L.u32p), L.ptr(lo, L.u32p), L.ptr(hi, L.u32p), idx.size,
                                           C.byref(h)))
        return cls(device, h, result.voxel_resolution)

    def _require(self):
        if not self._h:
            raise RuntimeError("VoxelGrid is closed")
        return L.load()

    @property
    def count(self) -> int:
        return int(self._require().dl_voxel_result_count(self._h))

    @property
    def device_ms(self) -> float:
        return float(self._require().dl_voxel_result_device_ms(self._h))

    def stats(self) -> dict:
        """Timing and sizes of the voxelize call that made this grid
        (dl_voxel_result_stats): device_ms, eval_ms (the per-voxel sums),
        block_pairs, grid_blocks."""
        st = L.DlVoxelStats()
        L.check(self._require().dl_voxel_result_stats(self._h, C.byref(st)))
        return {"device_ms": float(st.device_ms), "eval_ms": float(st.eval_ms),
                "block_pairs": int(st.block_pairs), "grid_blocks": int(st.grid_blocks)}

    def device_view(self) -> "L.DlVoxelDeviceView":
        v = L.DlVoxelDeviceView()
        L.check(self._require().dl_voxel_result_device(self._h, C.byref(v)))
        return v

    def result(self) -> VoxelizeResult:
        """The grid as a host VoxelizeResult (voxelize.ts:38-45)."""
        lib = self._require()
        n = int(lib.dl_voxel_result_count(self._h))
        idx = np.empty(n, np.uint32)
        lo = np.empty(n, np.uint32)
        hi = np.empty(n, np.uint32)
        L.check(lib.dl_voxel_result_blocks(self._h, L.ptr(idx, L.u32p), L.ptr(lo, L.u32p), L.ptr(hi, L.u32p)))
        bmin, bmax, nb = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        L.check(lib.dl_voxel_result_bounds(self._h, bmin, bmax, nb))
        return VoxelizeResult(BlockMaskBuffer(idx, lo, hi), Bounds(tuple(bmin), tuple(bmax)),
                              self.voxel_resolution, nb[0], nb[1], nb[2], self.device_ms)

    def fill(self) -> "VoxelGrid":
        """Closes every cavity that solid voxels enclose."""
        lib = self._require()
        h = C.c_void_p()
        L.check(lib.dl_voxel_fill(self._device.handle, self._h, C.byref(h)))
        return VoxelGrid(self._device, h, self.voxel_resolution)

    def carve(self, points) -> "VoxelGrid":
        """Makes solid every voxel that cannot be reached from the points."""
        lib = self._require()
        pts = _points(points)
        h = C.c_void_p()
        L.check(lib.dl_voxel_carve(self._device.handle, self._h, L.ptr(pts, L.f64p), pts.shape[0], C.byref(h)))
        return VoxelGrid(self._device, h, self.voxel_resolution)

    def query(self, points) -> np.ndarray:
        """Occupancy (bool) of the voxels holding the points; False outside."""
        lib = self._require()
        pts = _points(points)
        out = np.empty(pts.shape[0], np.uint8)
        L.check(lib.dl_voxel_query(self._device.handle, self._h, L.ptr(pts, L.f64p), pts.shape[0],
                                   L.ptr(out, L.u8p)))
        return out.astype(bool)

    def mesh(self) -> VoxelMesh:
        lib = self._require()
        m = C.c_void_p()
        L.check(lib.dl_voxel_mesh_build(self._device.handle, self._h, C.byref(m)))
        try:
            q = int(lib.dl_voxel_mesh_quads(m))
            pos = np.empty(q * 12, np.float32)
            idx = np.empty(q * 6, np.uint32)
            L.check(lib.dl_voxel_mesh_read(m, L.ptr(pos, L.f32p), L.ptr(idx, L.u32p)))
            ms = float(lib.dl_voxel_mesh_device_ms(m))
        finally:
            lib.dl_voxel_mesh_free(m)
        return VoxelMesh(pos.reshape(-1, 3), idx.reshape(-1, 3), ms)

    def close(self) -> None:
        if self._h:
            L.load().dl_voxel_result_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def voxelize_resident(device, data: SplatData, opts: Optional[VoxelizeOptions] = None) -> VoxelGrid:
    """`voxelize` whose result stays in HBM as a VoxelGrid."""
    opts = opts or VoxelizeOptions()
    o = _options_struct(opts)
    lib = L.load()
    h = C.c_void_p()
    s = L.splats_struct(data)
    L.check(lib.dl_voxelize(device.handle, C.byref(s), C.byref(o), C.byref(h)))
    return VoxelGrid(device, h, opts.voxel_resolution)


def voxelize(device, data: SplatData, opts: Optional[VoxelizeOptions] = None) -> VoxelizeResult:
    """`voxelize(device, data, opts)` (voxelize.ts:59-248)."""
    with voxelize_resident(device, data, opts) as grid:
        return grid.result()
