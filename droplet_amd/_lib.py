"""ctypes binding of libdroplet_hip.so (include/droplet_hip.h).

The library is loaded from this package directory only; when it is missing the
import of any device entry point raises, so no code path can silently fall
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; DROPLET_HIP_LIB names another build of the same library
LIB_PATH = os.environ.get("DROPLET_HIP_LIB") or os.path.join(HERE, "libdroplet_hip.so")

DL_OK = 0
DL_ERR_INVALID = 1
DL_ERR_DEVICE = 2
DL_ERR_OOM = 3
DL_ERR_CAPACITY = 4
DL_ERR_INTERNAL = 5

STAGES = ("cull", "depth_sort", "project", "scan", "emit", "pair_sort", "ranges", "rasterize")
TAPS = ("sorted_indices", "depth_keys", "projected", "tile_counts", "pair_tiles", "pair_splats",
        "tile_offsets", "accum")

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)


class DlSplats(C.Structure):
    _fields_ = [("count", C.c_uint32), ("positions", f32p), ("rotations", f32p), ("log_scales", f32p),
                ("opacity_logits", f32p), ("colors_dc", f32p), ("colors_sh", f32p), ("sh_bands", C.c_uint32)]


class DlCamera(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("view_matrix", C.c_float * 16),
                ("fov_y", C.c_double), ("near_plane", C.c_double), ("far_plane", C.c_double),
                ("projection", C.c_uint32)]


class DlBackground(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("a", C.c_float)]


class DlDeviceOpts(C.Structure):
    _fields_ = [("device_index", C.c_int)]


class DlRendererOpts(C.Structure):
    _fields_ = [("chunk_cap", C.c_uint32), ("max_coverage_per_splat", C.c_uint32), ("binning", C.c_uint32),
                ("first_pass", C.c_uint32), ("pass_growth", C.c_uint32), ("full_lists", C.c_uint32),
                ("sort_digit_bits", C.c_uint32), ("tile_size", C.c_uint32), ("raster_group", C.c_uint32)]


class DlRenderStats(C.Structure):
    _fields_ = [("num_visible", C.c_uint32), ("num_chunks", C.c_uint32), ("num_tiles", C.c_uint32),
                ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32), ("pair_key_bits", C.c_uint32),
                ("num_pairs", C.c_uint64), ("stage_ms", C.c_float * 8), ("total_ms", C.c_float),
                ("pixel_iters", C.c_uint64), ("pairs_consumed", C.c_uint64), ("list_entries", C.c_uint64),
                ("binned_splats", C.c_uint64), ("row_spans", C.c_uint64), ("group_spans", C.c_uint64),
                ("depth_key_bits", C.c_uint32), ("tile_size", C.c_uint32), ("raster_group", C.c_uint32),
                ("coverage_truncated", C.c_uint32)]


class DlGrads(C.Structure):
    _fields_ = [("positions", f32p), ("rotations", f32p), ("log_scales", f32p), ("opacity_logits", f32p),
                ("colors_dc", f32p), ("colors_sh", f32p)]


class DlTuneResult(C.Structure):
    _fields_ = [("chosen", DlRendererOpts), ("chosen_ms", C.c_float), ("default_ms", C.c_float),
                ("candidates", C.c_uint32)]


class DlView(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("camera", DlCamera), ("background", DlBackground)]


class DlBatchStats(C.Structure):
    _fields_ = [("num_views", C.c_uint32), ("num_groups", C.c_uint32), ("graph_replays", C.c_uint32),
                ("captures", C.c_uint32), ("num_pairs", C.c_uint64), ("device_ms", C.c_float)]


class DlVoxelOpts(C.Structure):
    _fields_ = [("voxel_resolution", C.c_double), ("opacity_cutoff", C.c_double), ("has_bounds", C.c_int),
                ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3)]


class DlVoxelDeviceView(C.Structure):
    _fields_ = [("masks", C.c_void_p), ("num_blocks", C.c_uint64), ("nb", C.c_int32 * 3),
                ("bounds_min", C.c_double * 3), ("voxel_resolution", C.c_double), ("block_idx", C.c_void_p),
                ("mask_lo", C.c_void_p), ("mask_hi", C.c_void_p), ("count", C.c_uint64), ("device", C.c_int)]


class DlVoxelStats(C.Structure):
    _fields_ = [("device_ms", C.c_float), ("eval_ms", C.c_float), ("block_pairs", C.c_uint64),
                ("grid_blocks", C.c_uint64)]


f64p = C.POINTER(C.c_double)

# (name, restype, argtypes) of every entry point declared in include/droplet_hip.h
SIGNATURES = [
    ("dl_context_create", C.c_int, [C.POINTER(DlDeviceOpts), C.POINTER(C.c_void_p)]),
    ("dl_context_destroy", None, [C.c_void_p]),
    ("dl_last_error", C.c_char_p, []),
    ("dl_context_device_name", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("dl_renderer_create", C.c_int, [C.c_void_p, C.POINTER(DlRendererOpts), C.POINTER(C.c_void_p)]),
    ("dl_renderer_destroy", None, [C.c_void_p]),
    ("dl_render", C.c_int, [C.c_void_p, C.POINTER(DlSplats), C.POINTER(DlCamera), C.POINTER(DlBackground), u8p]),
    ("dl_scene_create", C.c_int, [C.c_void_p, C.POINTER(DlSplats), C.POINTER(C.c_void_p)]),
    ("dl_scene_destroy", None, [C.c_void_p]),
    ("dl_scene_create_from_ply", C.c_int, [C.c_void_p, u8p, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    ("dl_scene_info", C.c_int, [C.c_void_p, u32p, u32p, u32p]),
    ("dl_scene_read", C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p, f32p, f32p]),
    ("dl_scene_bounds", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("dl_render_scene", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DlCamera), C.POINTER(DlBackground), u8p]),
    ("dl_renderer_read_image", C.c_int, [C.c_void_p, u8p]),
    ("dl_renderer_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("dl_renderer_set_counting", C.c_int, [C.c_void_p, C.c_int]),
    ("dl_renderer_stats", C.c_int, [C.c_void_p, C.POINTER(DlRenderStats)]),
    ("dl_renderer_row_iters", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint32, u32p]),
    ("dl_renderer_set_pair_budget", C.c_int, [C.c_void_p, C.c_uint64]),
    ("dl_renderer_set_taps", C.c_int, [C.c_void_p, C.c_int]),
    ("dl_renderer_tap", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("dl_render_forward", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DlCamera), C.POINTER(DlBackground), f32p]),
    ("dl_render_backward", C.c_int, [C.c_void_p, f32p, C.POINTER(DlGrads)]),
    ("dl_renderer_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DlCamera), C.POINTER(DlTuneResult)]),
    ("dl_renderer_tune_candidates", C.c_int, [C.POINTER(DlRendererOpts), C.c_uint64, C.POINTER(DlRendererOpts),
                                              C.c_uint32, C.POINTER(C.c_uint32)]),
    ("dl_batch_autotune", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(DlView), C.c_uint32,
                                    C.POINTER(DlTuneResult)]),
    ("dl_batch_create", C.c_int, [C.c_void_p, C.POINTER(DlRendererOpts), C.POINTER(C.c_void_p)]),
    ("dl_batch_destroy", None, [C.c_void_p]),
    ("dl_batch_set_graph", C.c_int, [C.c_void_p, C.c_int]),
    ("dl_batch_render", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(DlView), C.c_uint32]),
    ("dl_batch_set_pair_budget", C.c_int, [C.c_void_p, C.c_uint64]),
    ("dl_batch_read_image", C.c_int, [C.c_void_p, C.c_uint32, u8p]),
    ("dl_batch_get_stats", C.c_int, [C.c_void_p, C.POINTER(DlBatchStats)]),
    ("dl_sort_pairs_u32", C.c_int, [C.c_void_p, u32p, u32p, C.c_uint32, C.c_uint32]),
    ("dl_sort_pairs_u32_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("dl_sort_pairs_u32_noalloc_max", C.c_uint64, []),
    ("dl_scan_u32", C.c_int, [C.c_void_p, u32p, u32p, C.c_uint64, C.c_uint32, u32p]),
    ("dl_voxelize", C.c_int, [C.c_void_p, C.POINTER(DlSplats), C.POINTER(DlVoxelOpts), C.POINTER(C.c_void_p)]),
    ("dl_voxel_result_count", C.c_uint64, [C.c_void_p]),
    ("dl_voxel_result_blocks", C.c_int, [C.c_void_p, u32p, u32p, u32p]),
    ("dl_voxel_result_bounds", C.c_int, [C.c_void_p, C.c_double * 3, C.c_double * 3, C.c_double * 3]),
    ("dl_voxel_result_device_ms", C.c_float, [C.c_void_p]),
    ("dl_voxel_result_stats", C.c_int, [C.c_void_p, C.POINTER(DlVoxelStats)]),
    ("dl_voxel_result_free", None, [C.c_void_p]),
    ("dl_voxel_result_device", C.c_int, [C.c_void_p, C.POINTER(DlVoxelDeviceView)]),
    ("dl_voxel_result_create", C.c_int, [C.c_void_p, C.c_double * 3, C.c_double * 3, C.c_double * 3, C.c_double,
                                         u32p, u32p, u32p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("dl_voxel_fill", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("dl_voxel_carve", C.c_int, [C.c_void_p, C.c_void_p, f64p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("dl_voxel_query", C.c_int, [C.c_void_p, C.c_void_p, f64p, C.c_uint64, u8p]),
    ("dl_voxel_mesh_build", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("dl_voxel_mesh_quads", C.c_uint64, [C.c_void_p]),
    ("dl_voxel_mesh_read", C.c_int, [C.c_void_p, f32p, u32p]),
    ("dl_voxel_mesh_device_ms", C.c_float, [C.c_void_p]),
    ("dl_voxel_mesh_free", None, [C.c_void_p]),
    ("dl_debug_math", C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p, C.c_uint32]),
    ("dl_debug_valu_probe", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    ("dl_context_synchronize", C.c_int, [C.c_void_p]),
]

_lib = None
_lock = threading.Lock()


class DropletError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def load(path: str = LIB_PATH):
    """Load libdroplet_hip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise FileNotFoundError(
                    f"{path} is missing: build it with `python -m droplet_amd.build` "
                    "(there is no CPU fallback)")
            L = C.CDLL(path)
            for name, res, args in SIGNATURES:
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != DL_OK:
        msg = load().dl_last_error()
        raise DropletError(rc, msg.decode() if msg else f"error {rc}")


def ptr(a, t=f32p):
    return None if a is None else a.ctypes.data_as(t)


def splats_struct(s) -> DlSplats:
    o = DlSplats()
    o.count = s.count
    o.positions = ptr(s.positions)
    o.rotations = ptr(s.rotations)
    o.log_scales = ptr(s.log_scales)
    o.opacity_logits = ptr(s.opacity_logits)
    o.colors_dc = ptr(s.colors_dc)
    o.colors_sh = ptr(s.colors_sh)
    o.sh_bands = s.sh_bands
    return o


def camera_struct(cam) -> DlCamera:
    c = DlCamera()
    c.width, c.height = cam.width, cam.height
    vm = np.asarray(cam.view_matrix, dtype=np.float32)
    for k in range(16):
        c.view_matrix[k] = float(vm[k])
    c.fov_y, c.near_plane, c.far_plane = float(cam.fov_y), float(cam.near), float(cam.far)
    c.projection = 0 if cam.projection == "pinhole" else 1
    return c


def background_struct(bg) -> DlBackground:
    b = DlBackground()
    if bg is None:
        b.r, b.g, b.b, b.a = 0.0, 0.0, 0.0, 1.0
    else:
        b.r, b.g, b.b, b.a = bg.r, bg.g, bg.b, bg.a
    return b
