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


def voxelize(device, data: SplatData, opts: Optional[VoxelizeOptions] = None) -> VoxelizeResult:
    """`voxelize(device, data, opts)` (voxelize.ts:59-248)."""
    opts = opts or VoxelizeOptions()
    o = L.DlVoxelOpts()
    o.voxel_resolution = float(opts.voxel_resolution)
    o.opacity_cutoff = float(opts.opacity_cutoff)
    if opts.bounds is not None:
        o.has_bounds = 1
        for k in range(3):
            o.bounds_min[k] = float(opts.bounds.min[k])
            o.bounds_max[k] = float(opts.bounds.max[k])
    lib = L.load()
    h = C.c_void_p()
    s = L.splats_struct(data)
    L.check(lib.dl_voxelize(device.handle, C.byref(s), C.byref(o), C.byref(h)))
    try:
        n = int(lib.dl_voxel_result_count(h))
        idx = np.empty(n, np.uint32)
        lo = np.empty(n, np.uint32)
        hi = np.empty(n, np.uint32)
        L.check(lib.dl_voxel_result_blocks(h, L.ptr(idx, L.u32p), L.ptr(lo, L.u32p), L.ptr(hi, L.u32p)))
        bmin, bmax, nb = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        L.check(lib.dl_voxel_result_bounds(h, bmin, bmax, nb))
        ms = float(lib.dl_voxel_result_device_ms(h))
    finally:
        lib.dl_voxel_result_free(h)
    return VoxelizeResult(BlockMaskBuffer(idx, lo, hi), Bounds(tuple(bmin), tuple(bmax)),
                          float(opts.voxel_resolution), nb[0], nb[1], nb[2], ms)
