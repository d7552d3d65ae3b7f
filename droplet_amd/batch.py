"""Batched views over resident scenes (`dl_batch_*` in include/droplet_hip.h).

No reference counterpart: it stands beside `Renderer.render`
(src/render/pipeline.ts:181-185) and reproduces, bit for bit, the image that
`Renderer.render(data, camera, background)` gives for every (scene, camera)
view. One call runs every stage once over all views. The first call with a
new layout captures the launches in a hipGraph, and later calls replay it
with new cameras."""
from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .renderer import Device, RendererOptions, Scene, tune_result_dict
from .types import BackgroundRGBA, RenderCamera


@dataclass
class View:
    scene: int                          # index into the scenes list of render()
    camera: RenderCamera
    background: Optional[BackgroundRGBA] = None


class BatchRenderer:
    def __init__(self, device: Device, opts: Optional[RendererOptions] = None, use_graph: bool = True):
        self.device = device
        opts = opts or RendererOptions()
        o = L.DlRendererOpts()
        o.max_coverage_per_splat = int(opts.max_coverage_per_splat or 0)
        o.chunk_cap = int(opts.chunk_cap or 0)   # scene splats per group
        o.first_pass = int(opts.first_pass or 0)
        o.pass_growth = int(opts.pass_growth or 0)
        o.full_lists = 1 if opts.full_lists else 0
        o.sort_digit_bits = int(opts.sort_digit_bits or 0)
        h = C.c_void_p()
        L.check(L.load().dl_batch_create(device.handle, C.byref(o), C.byref(h)))
        self._h = h
        self._sizes: List[tuple] = []
        self.set_graph(use_graph)

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("batch renderer destroyed")
        return self._h

    def set_graph(self, enabled: bool) -> None:
        L.check(L.load().dl_batch_set_graph(self.handle, 1 if enabled else 0))

    def render(self, scenes: Sequence[Scene], views: Sequence[View]) -> None:
        """Renders every view; images stay on the device (image(k))."""
        sc = (C.c_void_p * len(scenes))(*[s._h for s in scenes])
        vs = self._views_struct(views)
        L.check(L.load().dl_batch_render(self.handle, sc, len(scenes), vs, len(views)))
        self._sizes = [(v.camera.width, v.camera.height) for v in views]

    def autotune(self, scenes: Sequence[Scene], views: Sequence[View]) -> dict:
        """Times candidate options on this batch and keeps the fastest
        (dl_batch_autotune); the images are those of the last call."""
        sc = (C.c_void_p * len(scenes))(*[s._h for s in scenes])
        vs = self._views_struct(views)
        res = L.DlTuneResult()
        L.check(L.load().dl_batch_autotune(self.handle, sc, len(scenes), vs, len(views), C.byref(res)))
        self._sizes = [(v.camera.width, v.camera.height) for v in views]
        return tune_result_dict(res)

    def set_pair_budget(self, pairs: int) -> None:
        """Groups above the budget split by views and a view above it alone
        renders through the single-view path (dl_batch_set_pair_budget);
        0 restores the default 2^31."""
        L.check(L.load().dl_batch_set_pair_budget(self.handle, int(pairs)))

    def image(self, k: int) -> np.ndarray:
        w, h = self._sizes[k]
        out = np.empty(w * h * 4, np.uint8)
        L.check(L.load().dl_batch_read_image(self.handle, k, L.ptr(out, L.u8p)))
        return out

    @staticmethod
    def _views_struct(views: Sequence[View]):
        vs = (L.DlView * len(views))()
        for k, v in enumerate(views):
            vs[k].scene = v.scene
            vs[k].camera = L.camera_struct(v.camera)
            vs[k].background = L.background_struct(v.background or BackgroundRGBA())
        return vs

    def stats(self) -> dict:
        st = L.DlBatchStats()
        L.check(L.load().dl_batch_get_stats(self.handle, C.byref(st)))
        return {"num_views": st.num_views, "num_groups": st.num_groups, "graph_replays": st.graph_replays,
                "captures": st.captures, "num_pairs": int(st.num_pairs), "device_ms": float(st.device_ms)}

    def destroy(self) -> None:
        if getattr(self, "_h", None) is not None:
            L.load().dl_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be going away: the
        # process releases the device memory by itself
        if sys.is_finalizing():
            return
        try:
            self.destroy()
        except Exception:
            pass
