"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
bench's CPU-baseline leg, always as the checker and never as the thing that
is measured or shipped. See oracle.h for what is restated from which
reference file.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)


class OrcSplats(C.Structure):
    _fields_ = [("count", C.c_uint32), ("positions", f32p), ("rotations", f32p), ("log_scales", f32p),
                ("opacity_logits", f32p), ("colors_dc", f32p), ("colors_sh", f32p), ("sh_bands", C.c_uint32)]


class OrcCamera(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("view", C.c_float * 16),
                ("fov_y", C.c_double), ("near_plane", C.c_double), ("far_plane", C.c_double),
                ("projection", C.c_uint32)]


class OrcTaps(C.Structure):
    _fields_ = [("num_visible", C.c_uint32), ("num_chunks", C.c_uint32), ("num_pairs", C.c_uint64),
                ("num_tiles", C.c_uint32), ("pixel_iters", C.c_uint64), ("pixel_evals", C.c_uint64),
                ("pixel_blends", C.c_uint64), ("tile_iters", C.c_uint64),
                ("sorted_indices", u32p), ("depth_keys", u32p), ("projected", f32p), ("tile_counts", u32p),
                ("pair_tiles", u32p), ("pair_splats", u32p), ("tile_offsets", u32p), ("chunk_pairs", u32p),
                ("accum", f32p), ("cap_visible", C.c_uint64), ("cap_pairs", C.c_uint64), ("cap_chunks", C.c_uint64)]


class OrcVoxelOpts(C.Structure):
    _fields_ = [("voxel_resolution", C.c_double), ("opacity_cutoff", C.c_double), ("has_bounds", C.c_int),
                ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3)]


class OrcVoxelResult(C.Structure):
    _fields_ = [("num_blocks", C.c_uint64), ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3),
                ("num_blocks_xyz", C.c_double * 3), ("num_batches", C.c_uint64), ("num_candidates", C.c_uint64),
                ("contrib_pairs", C.c_uint64), ("num_near", C.c_uint64), ("ref_tests", C.c_uint64),
                ("block_idx", u32p), ("mask_lo", u32p), ("mask_hi", u32p), ("cap_blocks", C.c_uint64),
                ("near_block", u32p), ("near_bit", u32p), ("near_alpha", f32p), ("cap_near", C.c_uint64),
                ("near_eps", C.c_double)]


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf_array.argtypes = [f32p, f32p, C.c_size_t]
        L.orc_fmaf_array.argtypes = [f32p, f32p, f32p, f32p, C.c_size_t]
        L.orc_float_to_sortable.restype = C.c_uint32
        L.orc_float_to_sortable.argtypes = [C.c_float]
        L.orc_u32sat.restype = C.c_uint32
        L.orc_u32sat.argtypes = [C.c_float]
        L.orc_focal.restype = C.c_double
        L.orc_focal.argtypes = [C.POINTER(OrcCamera)]
        L.orc_cull.restype = C.c_uint32
        L.orc_cull.argtypes = [f32p, C.c_uint32, C.POINTER(OrcCamera), u32p, f32p]
        L.orc_depth_sort.argtypes = [u32p, f32p, C.c_uint32]
        L.orc_project_packed.restype = C.c_uint32
        L.orc_project_packed.argtypes = [f32p, C.c_uint32, C.POINTER(OrcCamera), C.c_uint32, f32p]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [C.POINTER(OrcSplats), C.POINTER(OrcCamera), f32p, C.c_uint32, C.c_uint32,
                                 u8p, C.POINTER(OrcTaps), C.c_int]
        L.orc_render_rows.restype = C.c_int
        L.orc_render_rows.argtypes = [C.POINTER(OrcSplats), C.POINTER(OrcCamera), f32p, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, u8p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64), C.c_int]
        L.orc_render_timed.restype = C.c_int
        L.orc_render_timed.argtypes = [C.POINTER(OrcSplats), C.POINTER(OrcCamera), C.c_uint32, C.c_int,
                                       C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.orc_sort_pairs.argtypes = [u32p, u32p, C.c_uint32]
        L.orc_voxelize.restype = C.c_int
        L.orc_voxelize.argtypes = [C.POINTER(OrcSplats), C.POINTER(OrcVoxelOpts), C.POINTER(OrcVoxelResult), C.c_int]
        L.orc_aabb_extents.argtypes = [C.POINTER(OrcSplats), f32p]
        L.orc_bvh_order.restype = C.c_int
        L.orc_bvh_order.argtypes = [f32p, f32p, C.c_uint32, u32p]
        L.orc_morton_encode3.restype = C.c_uint32
        L.orc_morton_encode3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_morton_decode3.argtypes = [C.c_uint32, u32p]
        u64p, i64p, f64p = C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.POINTER(C.c_double)
        L.orc_vox_fill.restype = C.c_int
        L.orc_vox_fill.argtypes = [u64p, i64p, u64p]
        L.orc_vox_carve.restype = C.c_int
        L.orc_vox_carve.argtypes = [u64p, i64p, f64p, C.c_double, f64p, C.c_uint64, u64p]
        L.orc_vox_query.argtypes = [u64p, i64p, f64p, C.c_double, f64p, C.c_uint64, u8p]
        L.orc_vox_mesh.restype = C.c_uint64
        L.orc_vox_mesh.argtypes = [u64p, i64p, f64p, C.c_double, f32p, u32p, C.c_uint64]
        L.orc_gc_create.restype = C.c_void_p
        L.orc_gc_create.argtypes = [C.POINTER(OrcSplats), C.POINTER(OrcCamera), f32p, f32p, C.c_uint32, C.c_int]
        L.orc_gc_free.argtypes = [C.c_void_p]
        L.orc_gc_num_params.restype = C.c_uint32
        L.orc_gc_num_params.argtypes = [C.c_void_p]
        L.orc_gc_record.restype = C.c_int
        L.orc_gc_record.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]
        L.orc_gc_image.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.orc_gc_fd_record.restype = C.c_int
        L.orc_gc_fd_record.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, C.POINTER(C.c_double)]
        L.orc_gc_fd.restype = C.c_int
        L.orc_gc_fd.argtypes = [C.c_void_p, C.c_uint32, u32p, u32p, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(C.c_int32), C.c_int]
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray], t):
    return None if a is None else a.ctypes.data_as(t)


def _splats(s) -> OrcSplats:
    o = OrcSplats()
    o.count = s.count
    o.positions = _p(s.positions, f32p)
    o.rotations = _p(s.rotations, f32p)
    o.log_scales = _p(s.log_scales, f32p)
    o.opacity_logits = _p(s.opacity_logits, f32p)
    o.colors_dc = _p(s.colors_dc, f32p)
    o.colors_sh = _p(s.colors_sh, f32p)
    o.sh_bands = s.sh_bands
    return o


def camera_struct(cam) -> OrcCamera:
    c = OrcCamera()
    c.width, c.height = cam.width, cam.height
    for k in range(16):
        c.view[k] = float(cam.view_matrix[k])
    c.fov_y, c.near_plane, c.far_plane = float(cam.fov_y), float(cam.near), float(cam.far)
    c.projection = 0 if cam.projection == "pinhole" else 1
    return c


def cpu_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


def expf(x) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))
    out = np.empty_like(x)
    lib().orc_expf_array(_p(x, f32p), _p(out, f32p), x.size)
    return out


def fmaf(a, b, c) -> np.ndarray:
    """Element-wise correctly rounded f32 fused multiply-add (C fmaf),
    broadcast over the arguments."""
    a, b, c = np.broadcast_arrays(np.asarray(a, np.float32), np.asarray(b, np.float32), np.asarray(c, np.float32))
    shape = a.shape
    a, b, c = (np.ascontiguousarray(v.reshape(-1)) for v in (a, b, c))
    out = np.empty_like(a)
    lib().orc_fmaf_array(_p(a, f32p), _p(b, f32p), _p(c, f32p), _p(out, f32p), a.size)
    return out.reshape(shape)


def focal(cam) -> float:
    c = camera_struct(cam)
    return lib().orc_focal(C.byref(c))


def float_to_sortable(f: float) -> int:
    return lib().orc_float_to_sortable(C.c_float(f))


def u32sat(f: float) -> int:
    return lib().orc_u32sat(C.c_float(f))


def cull(splats, cam):
    n = splats.count
    idx = np.empty(max(n, 1), np.uint32)
    dep = np.empty(max(n, 1), np.float32)
    c = camera_struct(cam)
    k = lib().orc_cull(_p(splats.positions, f32p), n, C.byref(c), _p(idx, u32p), _p(dep, f32p))
    return idx[:k].copy(), dep[:k].copy()


def depth_sort(indices: np.ndarray, depths: np.ndarray) -> np.ndarray:
    idx = np.ascontiguousarray(indices, dtype=np.uint32).copy()
    dep = np.ascontiguousarray(depths, dtype=np.float32)
    lib().orc_depth_sort(_p(idx, u32p), _p(dep, f32p), idx.size)
    return idx


def project_packed(rec: np.ndarray, bands: int, cam, max_coverage: int = 0):
    rec = np.ascontiguousarray(rec, dtype=np.float32)
    out = np.zeros(12, np.float32)
    c = camera_struct(cam)
    cnt = lib().orc_project_packed(_p(rec, f32p), bands, C.byref(c), max_coverage, _p(out, f32p))
    return out, cnt


def sort_pairs(keys: np.ndarray, vals: np.ndarray):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    v = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    lib().orc_sort_pairs(_p(k, u32p), _p(v, u32p), k.size)
    return k, v


def render(splats, cam, background=None, chunk_cap: int = 0, max_coverage: int = 0, taps: bool = True,
           threads: Optional[int] = None, pair_cap: Optional[int] = None) -> dict:
    """Full reference pipeline. Returns rgba (H*W*4 uint8) and, with taps,
    the intermediate arrays; per-chunk arrays are concatenated."""
    L = lib()
    threads = threads or cpu_threads()
    s = _splats(splats)
    c = camera_struct(cam)
    bg = None
    if background is not None:
        bg = np.array([background.r, background.g, background.b, background.a], np.float32)
    rgba = np.zeros(cam.width * cam.height * 4, np.uint8)
    t = OrcTaps()
    out = {}
    if taps:
        n = max(splats.count, 1)
        tiles = ((cam.width + 15) // 16) * ((cam.height + 15) // 16)
        chunk = chunk_cap if chunk_cap else 200_000
        max_chunks = max(1, (n + chunk - 1) // chunk)
        if pair_cap is None:
            # first pass for sizes only
            probe = OrcTaps()
            L.orc_render(C.byref(s), C.byref(c), _p(bg, f32p), chunk_cap, max_coverage, None, C.byref(probe), threads)
            pair_cap = int(probe.num_pairs)
        arrays = {
            "sorted_indices": np.zeros(n, np.uint32), "depth_keys": np.zeros(n, np.uint32),
            "projected": np.zeros(12 * n, np.float32), "tile_counts": np.zeros(n, np.uint32),
            "pair_tiles": np.zeros(max(pair_cap, 1), np.uint32), "pair_splats": np.zeros(max(pair_cap, 1), np.uint32),
            "tile_offsets": np.zeros(max_chunks * (tiles + 1), np.uint32), "chunk_pairs": np.zeros(max_chunks, np.uint32),
            "accum": np.zeros(cam.width * cam.height * 4, np.float32),
        }
        for k, a in arrays.items():
            setattr(t, k, _p(a, f32p if a.dtype == np.float32 else u32p))
        t.cap_visible, t.cap_pairs, t.cap_chunks = n, max(pair_cap, 1), max_chunks
    rc = L.orc_render(C.byref(s), C.byref(c), _p(bg, f32p), chunk_cap, max_coverage, _p(rgba, u8p), C.byref(t), threads)
    if rc != 0:
        raise RuntimeError(f"orc_render failed ({rc})")
    out["rgba"] = rgba
    out.update(num_visible=t.num_visible, num_chunks=t.num_chunks, num_pairs=int(t.num_pairs), num_tiles=t.num_tiles,
               pixel_iters=int(t.pixel_iters), pixel_evals=int(t.pixel_evals), pixel_blends=int(t.pixel_blends),
               tile_iters=int(t.tile_iters))
    if taps:
        nv, npairs = t.num_visible, int(t.num_pairs)
        out["sorted_indices"] = arrays["sorted_indices"][:nv]
        out["depth_keys"] = arrays["depth_keys"][:nv]
        out["projected"] = arrays["projected"][:12 * nv].reshape(nv, 12)
        out["tile_counts"] = arrays["tile_counts"][:nv]
        out["pair_tiles"] = arrays["pair_tiles"][:npairs]
        out["pair_splats"] = arrays["pair_splats"][:npairs]
        out["tile_offsets"] = arrays["tile_offsets"][:t.num_chunks * (t.num_tiles + 1)].reshape(t.num_chunks, t.num_tiles + 1)
        out["chunk_pairs"] = arrays["chunk_pairs"][:t.num_chunks]
        out["accum"] = arrays["accum"].reshape(cam.height, cam.width, 4)
    return out


def render_rows(splats, cam, background=None, row_step: int = 16, row_phase: int = 0, chunk_cap: int = 0,
                max_coverage: int = 0, threads: Optional[int] = None) -> dict:
    """The reference pipeline on the tile rows ty % row_step == row_phase only
    (their pairs are emitted, sorted and rasterized; a tile's list is the
    same as in the full emit). Returns rgba (H*W*4, other rows hold the
    background), the kept tile rows, num_visible, num_pairs (every pair of
    the frame) and row_iters (loop iterations per tile row, 0 for rows not
    kept). chunk_cap 0 is the reference's 200000."""
    L = lib()
    threads = threads or cpu_threads()
    s = _splats(splats)
    c = camera_struct(cam)
    bg = None
    if background is not None:
        bg = np.array([background.r, background.g, background.b, background.a], np.float32)
    rgba = np.zeros(cam.width * cam.height * 4, np.uint8)
    tiles_y = (cam.height + 15) // 16
    row_iters = np.zeros(tiles_y, np.uint64)
    nvis, npairs = C.c_uint32(), C.c_uint64()
    rc = L.orc_render_rows(C.byref(s), C.byref(c), _p(bg, f32p), chunk_cap, max_coverage, row_step, row_phase,
                           _p(rgba, u8p), C.byref(nvis), C.byref(npairs),
                           row_iters.ctypes.data_as(C.POINTER(C.c_uint64)), threads)
    if rc != 0:
        raise RuntimeError(f"orc_render_rows failed ({rc})")
    rows = np.arange(row_phase % row_step, tiles_y, row_step)
    return {"rgba": rgba, "tile_rows": rows, "num_visible": int(nvis.value), "num_pairs": int(npairs.value),
            "row_iters": row_iters}


def render_timed(splats, cam, row_step: int = 1, threads: Optional[int] = None):
    L = lib()
    threads = threads or cpu_threads()
    s = _splats(splats)
    c = camera_struct(cam)
    secs = (C.c_double * 7)()
    npairs = C.c_uint64()
    rc = L.orc_render_timed(C.byref(s), C.byref(c), row_step, threads, secs, C.byref(npairs))
    if rc != 0:
        raise RuntimeError("orc_render_timed failed")
    return list(secs), int(npairs.value)


def aabb_extents(splats) -> np.ndarray:
    out = np.zeros(max(3 * splats.count, 1), np.float32)
    s = _splats(splats)
    lib().orc_aabb_extents(C.byref(s), _p(out, f32p))
    return out[:3 * splats.count]


def bvh_order(positions: np.ndarray, extents: np.ndarray) -> np.ndarray:
    n = positions.size // 3
    out = np.zeros(max(n, 1), np.uint32)
    rc = lib().orc_bvh_order(_p(np.ascontiguousarray(positions, np.float32), f32p),
                             _p(np.ascontiguousarray(extents, np.float32), f32p), n, _p(out, u32p))
    if rc != 0:
        raise RuntimeError("bvh order incomplete")
    return out[:n]


def voxelize(splats, voxel_resolution: float = 0.05, opacity_cutoff: float = 0.1, bounds=None,
             near_eps: float = 1e-6, threads: Optional[int] = None, block_cap: Optional[int] = None,
             near_cap: int = 1 << 20) -> dict:
    L = lib()
    threads = threads or cpu_threads()
    s = _splats(splats)
    o = OrcVoxelOpts()
    o.voxel_resolution, o.opacity_cutoff = voxel_resolution, opacity_cutoff
    if bounds is not None:
        o.has_bounds = 1
        for k in range(3):
            o.bounds_min[k], o.bounds_max[k] = bounds.min[k], bounds.max[k]
    r = OrcVoxelResult()
    if block_cap is None:
        block_cap = 1 << 22
    arrs = {"block_idx": np.zeros(block_cap, np.uint32), "mask_lo": np.zeros(block_cap, np.uint32),
            "mask_hi": np.zeros(block_cap, np.uint32), "near_block": np.zeros(near_cap, np.uint32),
            "near_bit": np.zeros(near_cap, np.uint32), "near_alpha": np.zeros(near_cap, np.float32)}
    for k, a in arrs.items():
        setattr(r, k, _p(a, f32p if a.dtype == np.float32 else u32p))
    r.cap_blocks, r.cap_near, r.near_eps = block_cap, near_cap, near_eps
    rc = L.orc_voxelize(C.byref(s), C.byref(o), C.byref(r), threads)
    if rc != 0:
        raise RuntimeError(f"orc_voxelize failed ({rc})")
    nb, nn = int(r.num_blocks), int(r.num_near)
    if nb > block_cap:
        raise RuntimeError("block capacity too small")
    return {
        "block_idx": arrs["block_idx"][:nb].copy(), "mask_lo": arrs["mask_lo"][:nb].copy(),
        "mask_hi": arrs["mask_hi"][:nb].copy(),
        "bounds_min": tuple(r.bounds_min), "bounds_max": tuple(r.bounds_max),
        "num_blocks_xyz": tuple(r.num_blocks_xyz), "num_batches": int(r.num_batches),
        "num_candidates": int(r.num_candidates), "contrib_pairs": int(r.contrib_pairs),
        "ref_tests": int(r.ref_tests),
        "num_near": nn, "near_block": arrs["near_block"][:min(nn, near_cap)].copy(),
        "near_bit": arrs["near_bit"][:min(nn, near_cap)].copy(),
        "near_alpha": arrs["near_alpha"][:min(nn, near_cap)].copy(),
    }


def block_insertion_order(nb) -> np.ndarray:
    """Flat indices of every block of an nb[0] x nb[1] x nb[2] grid in the order
    voxelize adds them to its BlockMaskBuffer: 16^3-block batches in (z, y, x)
    loop order (voxelize.ts:186-195), and inside a batch bi = lx + ly*cx +
    lz*cx*cy (voxelize.ts:166-176). Empty blocks are skipped by addBlock, so
    the non-empty blocks of any grid appear in this order."""
    nbx, nby, nbz = (int(v) for v in nb)
    out = []
    for bz in range(0, nbz, 16):
        for by in range(0, nby, 16):
            for bx in range(0, nbx, 16):
                cx, cy, cz = min(16, nbx - bx), min(16, nby - by), min(16, nbz - bz)
                lz, ly, lx = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
                flat = (bx + lx) + (by + ly) * nbx + (bz + lz) * nbx * nby
                out.append(flat.ravel())
    return np.concatenate(out).astype(np.int64) if out else np.zeros(0, np.int64)


def sparse_blocks(dense: np.ndarray, nb):
    """(block_idx, mask_lo, mask_hi) of the non-empty blocks of a dense grid in
    voxelize's insertion order."""
    order = block_insertion_order(nb)
    m = np.asarray(dense, np.uint64)[order]
    keep = m != 0
    return (order[keep].astype(np.uint32), (m[keep] & np.uint64(0xFFFFFFFF)).astype(np.uint32),
            (m[keep] >> np.uint64(32)).astype(np.uint32))


def dense_masks(block_idx, mask_lo, mask_hi, nb) -> np.ndarray:
    """Dense u64 grid (flat block index) of a BlockMaskBuffer's blocks; repeated
    blocks OR-merge as BlockMaskBuffer.addBlock does."""
    total = int(nb[0]) * int(nb[1]) * int(nb[2])
    out = np.zeros(total, np.uint64)
    m = np.asarray(mask_lo, np.uint64) | (np.asarray(mask_hi, np.uint64) << np.uint64(32))
    np.bitwise_or.at(out, np.asarray(block_idx, np.int64), m)
    return out


def _grid_args(mask, nb):
    mask = np.ascontiguousarray(mask, np.uint64)
    nbv = np.ascontiguousarray(nb, np.int64)
    assert nbv.size == 3 and mask.size == int(nbv.prod())
    return mask, nbv


def vox_fill(mask, nb) -> np.ndarray:
    """Grid with every cavity enclosed by solid voxels closed (collision.c)."""
    mask, nbv = _grid_args(mask, nb)
    out = np.empty_like(mask)
    rc = lib().orc_vox_fill(_p(mask, C.POINTER(C.c_uint64)), _p(nbv, C.POINTER(C.c_int64)),
                            _p(out, C.POINTER(C.c_uint64)))
    assert rc == 0
    return out


def vox_carve(mask, nb, bmin, res: float, points) -> np.ndarray:
    """Grid where every voxel not reachable from the points is solid (collision.c)."""
    mask, nbv = _grid_args(mask, nb)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    bm = np.ascontiguousarray(bmin, np.float64)
    out = np.empty_like(mask)
    rc = lib().orc_vox_carve(_p(mask, C.POINTER(C.c_uint64)), _p(nbv, C.POINTER(C.c_int64)),
                             _p(bm, C.POINTER(C.c_double)), float(res), _p(pts, C.POINTER(C.c_double)),
                             pts.shape[0], _p(out, C.POINTER(C.c_uint64)))
    assert rc == 0
    return out


def vox_query(mask, nb, bmin, res: float, points) -> np.ndarray:
    mask, nbv = _grid_args(mask, nb)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    bm = np.ascontiguousarray(bmin, np.float64)
    out = np.empty(pts.shape[0], np.uint8)
    lib().orc_vox_query(_p(mask, C.POINTER(C.c_uint64)), _p(nbv, C.POINTER(C.c_int64)),
                        _p(bm, C.POINTER(C.c_double)), float(res), _p(pts, C.POINTER(C.c_double)),
                        pts.shape[0], _p(out, u8p))
    return out


def vox_mesh(mask, nb, bmin, res: float):
    """(positions float32 [Q*4, 3], indices uint32 [Q*2, 3]) of the quad mesh."""
    mask, nbv = _grid_args(mask, nb)
    bm = np.ascontiguousarray(bmin, np.float64)
    L = lib()
    args = (_p(mask, C.POINTER(C.c_uint64)), _p(nbv, C.POINTER(C.c_int64)), _p(bm, C.POINTER(C.c_double)),
            float(res))
    q = int(L.orc_vox_mesh(*args, None, None, 0))
    pos = np.empty(q * 12, np.float32)
    idx = np.empty(q * 6, np.uint32)
    q2 = int(L.orc_vox_mesh(*args, _p(pos, f32p), _p(idx, u32p), q))
    assert q2 == q
    return pos.reshape(-1, 3), idx.reshape(-1, 3)


def morton_encode3(x: int, y: int, z: int) -> int:
    return lib().orc_morton_encode3(x, y, z)


def morton_decode3(m: int):
    o = (C.c_uint32 * 3)()
    lib().orc_morton_decode3(m, o)
    return tuple(o)


# ---------------------------------------------------------------------------
# FP64 central differences of the differentiable render (gradcheck.c)
# ---------------------------------------------------------------------------
GC_NOT_VISIBLE, GC_RADIUS, GC_ORDER, GC_CULL, GC_BRANCH, GC_PROJ_KINK = 1, 2, 4, 8, 16, 32
# parameter index ranges of one splat (then SH, channel-major)
GC_POS, GC_ROT, GC_LOGSCALE, GC_LOGIT, GC_DC, GC_SH = 0, 3, 7, 10, 11, 14


class GradCheck:
    """L = sum (clamp(C + T*bg, 0, 1) - target)^2 over RGB, evaluated in f64
    with the f32 pipeline's lists; central differences per parameter."""

    def __init__(self, splats, cam, target_rgb: np.ndarray, background=None, max_coverage: int = 0,
                 threads: Optional[int] = None):
        self._keep = (splats, np.ascontiguousarray(target_rgb, dtype=np.float32).reshape(-1))
        self.threads = threads or min(16, cpu_threads())
        s = _splats(splats)
        c = camera_struct(cam)
        bg = None if background is None else np.array([background.r, background.g, background.b, background.a],
                                                        np.float32)
        self._bg = bg
        self._h = lib().orc_gc_create(C.byref(s), C.byref(c), _p(bg, f32p), _p(self._keep[1], f32p),
                                      max_coverage, self.threads)
        if not self._h:
            raise ValueError("gradient check needs a pinhole camera and shBands <= 3")
        self.width, self.height = cam.width, cam.height
        self.num_params = lib().orc_gc_num_params(self._h)

    def image(self):
        out = np.empty(self.width * self.height * 3, np.float64)
        loss = C.c_double()
        lib().orc_gc_image(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(loss), self.threads)
        return out.reshape(self.height, self.width, 3), loss.value

    def record(self, splat: int):
        out = np.empty(10, np.float64)
        rc = lib().orc_gc_record(self._h, splat, out.ctypes.data_as(C.POINTER(C.c_double)))
        return None if rc else out

    def fd(self, ids, params, h: float = 1e-4):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        params = np.ascontiguousarray(params, dtype=np.uint32)
        out = np.zeros(ids.size, np.float64)
        flags = np.zeros(ids.size, np.int32)
        lib().orc_gc_fd(self._h, ids.size, _p(ids, u32p), _p(params, u32p), h,
                        out.ctypes.data_as(C.POINTER(C.c_double)), flags.ctypes.data_as(C.POINTER(C.c_int32)),
                        self.threads)
        return out, flags

    def fd_record(self, splat: int, field: int, h: float = 1e-6):
        out = C.c_double()
        flags = lib().orc_gc_fd_record(self._h, splat, field, h, C.byref(out))
        return out.value, flags

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_gc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
