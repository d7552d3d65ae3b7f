"""Benchmark of the forward splat renderer on one MI355X.

Workload: 6,000,000 Gaussians (positions ~N(0, 0.75^2) clipped to [-3, 3],
SH bands 3), 3840x2160 pinhole from (0,0,4), scene resident in HBM. A step is
one full frame: f64 cull + depth keys + EWA projection, depth sort, colours,
binning into per-tile lists, rasterize + finalize to RGBA8 in device memory.
The RGBA8 readback is timed separately and is not part of `value`.

Prints one JSON line (metric/value in ms per frame, lower is better): `value`
is the median of the K timed frames, with p10 / p90 beside it and
`ms_per_step` the mean over the bracketed K frames. Also:
  roofline      dominant kernel (rasterize) against the MI355X FP32 peak,
                duration from hipEvents on the renderer's stream
  stages        per-stage hipEvent times with their algorithmic HBM bytes
                and fraction of the 8 TB/s peak
  equirect      the same scene at 4096x2048 equirect from the origin
  cpu_baseline  the CPU oracle on the same scene: every stage on all 6M
                splats, rasterize on every ROW_STEP-th tile row scaled up
  batched       512 views (8 scenes x 64 orbit cameras) as one batched call
                against the same views rendered one by one
  voxelize      the 1M-Gaussian voxelizer workload: device time, wall time
                with the host transfers, the k_vox_eval roofline and the
                oracle's CPU time on the same workload
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame @6M splats 4K; batched frames/s vs one-by-one; % HBM/VALU peak"
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (published; counts packed FP32 and FMA as two)
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E (published)
# Algorithmic work per rasterize loop iteration: the reference's per-pixel
# body (rasterize.ts:62-79) is 31 f32 operations counting each add, mul,
# compare, min/max and the exp as one.
RAST_OPS_PER_ITER = 31
# Voxelizer work per (voxel, Gaussian) test that passes the AABB test:
# ~55 f32 operations + 1 exp (voxel/shaders/voxelize.ts:62-84, exp(-ls) and
# the sigmoid hoisted per Gaussian); a rejected test is ~9 operations.
VOX_OPS_PER_PAIR = 56
VOX_OPS_PER_REJECT = 9
ROW_STEP = 16                # CPU baseline: rasterize every 16th tile row
# The probe of known VALU work that calibrates the SQ counters: blocks of
# 256 threads, each lane 16 * iters f32 FMAs
VALU_PROBE_BLOCKS = 2048
VALU_PROBE_ITERS = 4096


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _load_traffic():
    path = os.path.join(ROOT, "profiles", "rasterize_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _pct(values, q):
    return float(np.percentile(np.asarray(values, np.float64), q))


def cpu_baseline(data, cam):
    import oracle  # CPU checker only; never the measured product

    threads = min(16, oracle.cpu_threads())
    t0 = time.perf_counter()
    secs, npairs = oracle.render_timed(data, cam, row_step=ROW_STEP, threads=threads)
    wall = time.perf_counter() - t0
    # stage 6 holds rasterize over 1/ROW_STEP of the tile rows (+ nothing else)
    frame_s = sum(secs[:6]) + secs[6] * ROW_STEP
    return {
        "value": round(frame_s * 1e3, 1), "unit": "ms/frame", "cores": threads, "kind": "port",
        "sample": (f"C oracle on the same 6M-splat 4K frame: cull, depth sort, project, scan+emit, "
                   f"counting pair sort ({npairs} pairs) and ranges on the whole scene; rasterize on every "
                   f"{ROW_STEP}th tile row, scaled x{ROW_STEP}; {wall:.1f} s of CPU wall time; "
                   f"{_cpu_model()}"),
        "stage_s": [round(s, 3) for s in secs],
    }


def stage_table(stage_ms, st, n, bands):
    """Per-stage hipEvent times with their algorithmic HBM bytes and the
    fraction of the 8 TB/s peak those bytes take at that time. Stage names
    say what runs on the span-binning path; the byte figures are the
    minimum each stage must move (counted from a tally frame's own pass
    sizes at the timed tile size), not the bytes the counters measure."""
    passes = max(1, st["num_chunks"])
    items, spans, gspans = st["binned_splats"], st["row_spans"], st["group_spans"]
    entries = st["list_entries"]
    kbits = st["depth_key_bits"] or 32
    digit_passes = -(-kbits // 9) if kbits < 32 else 4
    color_rec = {0: 16, 1: 64, 2: 128, 3: 192}[bands]
    rows = [
        # f64 cull, depth key, EWA projection, packed rect: reads position,
        # rotation, log-scales, opacity (44 B), writes key, record, rect (60 B)
        ("cull_project", "k_cull_project + k_reduce_totals", 104 * n),
        # LSD passes: each reads the keys for the digit counts, then reads
        # key + value (the first pass's values are the indices) and writes both
        ("depth_sort", f"{digit_passes} radix passes", n * (4 * digit_passes + 12 + 16 * (digit_passes - 1))),
        # colours of the splats that reach a pass's lists: id, projection
        # record and camera position in, the colour record, colour out
        ("colors", "k_rects_colors / k_splat_colors", items * (4 + 16 + 12 + color_rec + 12)),
        # span binning: rects read twice, row spans and group spans written
        # once and read twice, one u32 per list entry, live bits of later passes
        ("binning", "span_bin (level 1, 2a, 2b, live compaction)",
         32 * items + 24 * spans + 24 * gspans + 4 * entries + 8 * n * (passes - 1)),
    ]
    out = {}
    names = {"cull_project": "cull", "depth_sort": "depth_sort", "colors": "project", "binning": "emit"}
    for name, kernels, nbytes in rows:
        ms = stage_ms.get(names[name], 0.0)
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[name] = {"ms": round(ms, 4), "kernels": kernels, "algorithmic_bytes": int(nbytes),
                     "GB_s": round(gbs, 1) if gbs else None,
                     "hbm_frac": round(gbs / PEAK_HBM_GBS, 4) if gbs else None}
    out["rasterize"] = {"ms": round(stage_ms.get("rasterize", 0.0), 4), "kernels": "rasterizer (per pass)",
                        "bound": "valu (see roofline)"}
    for k in ("scan", "pair_sort", "ranges"):  # the emit + pair sort path only
        if stage_ms.get(k, 0.0) > 0:
            out[k] = {"ms": round(stage_ms[k], 4)}
    return out


def time_frames(r, scene, cam, steps, dev):
    """K frames as a caller runs them (no stage events), each timed on the
    host clock (a frame returns once done); the K frames are bracketed by
    device synchronisations."""
    dev.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        r.render_scene(scene, cam, readback=False)
        per.append((time.perf_counter() - t1) * 1e3)
    dev.synchronize()
    return per, (time.perf_counter() - t0) * 1e3


def equirect_leg(dev, scene, note, steps=10):
    """The large scene at 4096x2048 equirect from the origin: 22.4e9 tile
    pairs, which no single pass's u32 lists hold, so it renders in passes
    cut from the per-splat tile counts."""
    from droplet_amd import synthetic as S
    from droplet_amd.renderer import Renderer

    cam = S.equirect_camera(4096, 2048)
    r = Renderer(dev)
    try:
        for _ in range(2):
            r.render_scene(scene, cam, readback=False)
        per, total = time_frames(r, scene, cam, steps, dev)
        st = r.stats()
    finally:
        r.destroy()
    note(f"equirect 4096x2048: median {np.median(per):.3f} ms/frame, {st['num_chunks']} passes, "
         f"{st['num_pairs']} pairs")
    return {"workload": "6M Gaussians, SH bands 3, 4096x2048 equirect from the origin, scene resident in HBM",
            "ms_per_frame": round(float(np.median(per)), 3), "p10_ms": round(_pct(per, 10), 3),
            "p90_ms": round(_pct(per, 90), 3), "frames": steps, "passes": st["num_chunks"],
            "pairs": st["num_pairs"], "list_entries": st["list_entries"], "visible": st["num_visible"]}


def batched_views(dev, steps: int, note, tune: bool = True):
    """8 resident scenes (50k .. 4M Gaussians, render distributions) x 64
    orbit cameras at 640x360 / 1280x720 / 1920x1080: 512 views. Plain
    one-by-one renders of every view against one batched call (graph
    replay); a sample of views is checked bit-exact between the two."""
    from droplet_amd import synthetic as S
    from droplet_amd.batch import BatchRenderer
    from droplet_amd.renderer import Renderer, Scene

    datas, views = S.batch_workload()
    scenes = [Scene(dev, d) for d in datas]
    del datas
    r = Renderer(dev)
    for v in views[::37]:
        r.render_scene(scenes[v.scene], v.camera, readback=False)
    dev.synchronize()
    t0 = time.perf_counter()
    for v in views:
        r.render_scene(scenes[v.scene], v.camera, readback=False)
    dev.synchronize()
    plain_s = time.perf_counter() - t0
    note(f"batched workload: one-by-one {plain_s * 1e3:.1f} ms for {len(views)} views")
    b = BatchRenderer(dev)
    tuned = None
    if tune:
        tuned = b.autotune(scenes, views)
        note(f"batched autotune: {tuned['chosen']} {tuned['chosen_ms']:.1f} ms per call "
             f"(default {tuned['default_ms']:.1f} ms)")
    b.render(scenes, views)   # plans, renders directly, captures the graph
    b.render(scenes, views)   # first replay
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render(scenes, views)
    dev.synchronize()
    batch_s = (time.perf_counter() - t0) / steps
    st = b.stats()
    checked = 0
    for k in range(0, len(views), 29):
        v = views[k]
        ref = r.render_scene(scenes[v.scene], v.camera)
        if not np.array_equal(b.image(k), ref):
            raise RuntimeError(f"batched view {k} differs from the single-view render")
        checked += 1
    note(f"batched workload: {batch_s * 1e3:.1f} ms per call ({st['num_groups']} groups, "
         f"{st['graph_replays']} graph replays), {checked} views verified bit-exact")
    out = {
        "workload": "8 scenes (50k-4M Gaussians) x 64 orbit cameras, 640x360/1280x720/1920x1080",
        "views": len(views), "scene_splats": [s.count for s in scenes],
        "frames_per_s": round(len(views) / batch_s, 1),
        "one_by_one_frames_per_s": round(len(views) / plain_s, 1),
        "speedup": round(plain_s / batch_s, 3),
        "ms_per_call": round(batch_s * 1e3, 3), "device_ms_per_call": round(st["device_ms"], 3),
        "groups": st["num_groups"], "pairs": st["num_pairs"], "views_verified_bitexact": checked,
        "autotune": tuned,
    }
    b.destroy()
    r.destroy()
    for s in scenes:
        s.destroy()
    return out


def voxelize_leg(dev, note, with_cpu: bool):
    """The voxelizer workload: 1M Gaussians uniform in a 10 x 3 x 10 m box,
    voxelResolution 0.02, opacityCutoff 0.1 (host arrays in, block list out,
    as voxelize() takes and returns them). The k_vox_eval roofline counts
    the reference's contributing (voxel, Gaussian) pairs, which the oracle
    counts on the same workload; the oracle's run is also the CPU baseline."""
    from droplet_amd import synthetic as S
    from droplet_amd.voxel import VoxelizeOptions, voxelize, voxelize_resident

    data, o = S.voxel_workload()
    opts = VoxelizeOptions(voxel_resolution=o["voxel_resolution"], opacity_cutoff=o["opacity_cutoff"])
    voxelize(dev, data, opts)  # warm
    walls, devs, evals = [], [], []
    for _ in range(3):
        t0 = time.perf_counter()
        with voxelize_resident(dev, data, opts) as g:
            res = g.result()
            vst = g.stats()
        walls.append((time.perf_counter() - t0) * 1e3)
        devs.append(vst["device_ms"])
        evals.append(vst["eval_ms"])
    note(f"voxelize: {min(devs):.3f} ms device ({min(evals):.3f} ms k_vox_eval), {min(walls):.3f} ms wall, "
         f"{res.buffer.count} blocks")
    out = {"workload": "1M Gaussians in a 10x3x10 m box, voxelResolution 0.02, opacityCutoff 0.1",
           "device_ms": round(min(devs), 3), "eval_ms": round(min(evals), 3),
           "wall_ms_incl_pcie": round(min(walls), 3), "blocks": int(res.buffer.count),
           "solid_voxels": int(res.buffer.popcount()), "block_pairs": vst["block_pairs"],
           "gaussians_per_s": round(data.count / (min(devs) * 1e-3), 1),
           "collision": collision_leg(dev, data, opts, note)}
    if with_cpu:
        import oracle  # CPU checker only; never the measured product

        threads = min(16, oracle.cpu_threads())
        t0 = time.perf_counter()
        ref = oracle.voxelize(data, voxel_resolution=opts.voxel_resolution, opacity_cutoff=opts.opacity_cutoff,
                              threads=threads, near_eps=0.0)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        if ref["block_idx"].size != res.buffer.count:
            raise RuntimeError("voxelizer block count differs from the oracle's")
        contrib, tests = ref["contrib_pairs"], ref["ref_tests"]
        eval_s = min(evals) * 1e-3
        achieved = contrib * VOX_OPS_PER_PAIR / eval_s / 1e12
        ref_rate = (contrib * VOX_OPS_PER_PAIR + (tests - contrib) * VOX_OPS_PER_REJECT) / eval_s / 1e12
        out["roofline"] = {
            "kernel": "k_vox_eval", "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "work": f"{contrib} contributing (voxel, Gaussian) pairs x {VOX_OPS_PER_PAIR} ops (counted by the oracle)",
            "reference_work": (f"{tests} AABB tests of the reference ({tests - contrib} rejected x "
                               f"{VOX_OPS_PER_REJECT} ops) at {ref_rate:.2f} TFLOP/s; the device's per-block "
                               f"lists skip almost all rejected tests, so this rate counts work it does not do"),
            "eval_ms": round(min(evals), 4),
        }
        out["cpu_baseline"] = {
            "value": round(cpu_ms, 1), "unit": "ms per voxelize", "cores": threads, "kind": "port",
            "sample": (f"C oracle on the same 1M-Gaussian workload, whole grid: f64 AABB extents, the BVH and its "
                       f"queries single-threaded (the reference's host passes), the per-voxel sums on {threads} "
                       f"OpenMP threads; {_cpu_model()}"),
        }
        note(f"voxelize roofline: {achieved:.2f} TFLOP/s ({contrib} contributing pairs in {min(evals):.3f} ms); "
             f"oracle {cpu_ms:.0f} ms on {threads} threads")
    return out


def collision_leg(dev, data, opts, note):
    """The stages after the voxelizer on its resident grid: fill, carve from
    the grid's corner voxel, and the quad mesh of the filled grid (stage
    times from hipEvents around each call, best of 3)."""
    from droplet_amd.voxel import voxelize_resident

    out = {}
    with voxelize_resident(dev, data, opts) as g:
        bmin = g.result().bounds.min
        seed = [[bmin[0] + 0.5 * opts.voxel_resolution, bmin[1] + 0.5 * opts.voxel_resolution,
                 bmin[2] + 0.5 * opts.voxel_resolution]]
        fills, carves, meshes = [], [], []
        for _ in range(3):
            with g.fill() as f:
                fills.append(f.device_ms)
                m = f.mesh()
                meshes.append(m.device_ms)
                filled = f.result().buffer.popcount()
            with g.carve(seed) as c:
                carves.append(c.device_ms)
                carved = c.result().buffer.popcount()
        out = {"fill_ms": round(min(fills), 3), "carve_ms": round(min(carves), 3),
               "mesh_ms": round(min(meshes), 3), "mesh_quads": int(m.quads),
               "filled_solid_voxels": int(filled), "carved_solid_voxels": int(carved)}
    note(f"collision: fill {out['fill_ms']} ms, carve {out['carve_ms']} ms, mesh {out['mesh_ms']} ms "
         f"({out['mesh_quads']} quads)")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--splats", type=int, default=6_000_000)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-views leg")
    ap.add_argument("--no-tune", action="store_true", help="time the default options (no autotune)")
    ap.add_argument("--tile-size", type=int, default=None, help="renderer tile size (8, 16, 32) before tuning")
    ap.add_argument("--raster-group", type=int, default=None, help="rasterizer group (1, 2) before tuning")
    ap.add_argument("--no-voxel", action="store_true", help="skip the voxelizer leg")
    ap.add_argument("--no-equirect", action="store_true", help="skip the equirect leg")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB device-to-device copy (a kernel of known traffic) so that a "
                         "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE run can calibrate the counters")
    ap.add_argument("--valu-probe", action="store_true",
                    help="also run one launch of k_valu_probe (a kernel of known VALU work) so that a "
                         "rocprofv3 --pmc SQ_INSTS_VALU ... run can calibrate the counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the renderer runs on a single "
              "device and multi-GPU sharding is not part of this build.", file=sys.stderr)
        return 2

    from droplet_amd import synthetic as S
    from droplet_amd import _lib as L
    from droplet_amd.renderer import Renderer, RendererOptions, Scene, create_device

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.copy_probe:
        import torch  # only the calibration copy uses torch

        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        note("copy probe: 1 GiB device-to-device copy done")
        del src, dst

    note(f"generating {args.splats} splats")
    data, cam = S.large_render_workload(n=args.splats)
    dev = create_device()
    if args.valu_probe:
        import ctypes as C

        chk = C.c_double()
        L.check(L.load().dl_debug_valu_probe(dev.handle, VALU_PROBE_BLOCKS, VALU_PROBE_ITERS, C.byref(chk)))
        note(f"VALU probe: {VALU_PROBE_BLOCKS} x 256 threads x {16 * VALU_PROBE_ITERS} FMAs (checksum {chk.value:.6g})")
    scene = Scene(dev, data)
    r = Renderer(dev, RendererOptions(tile_size=args.tile_size, raster_group=args.raster_group))
    note(f"scene resident on {dev.name}")

    tune = None
    if not args.no_tune:
        tune = r.autotune(scene, cam)
        note(f"autotune: {tune['chosen']} {tune['chosen_ms']:.3f} ms (default {tune['default_ms']:.3f} ms, "
             f"{tune['candidates']} candidates)")
    for _ in range(args.warmup):
        r.render_scene(scene, cam, readback=False)

    # the timed frames run as a caller runs them, without stage events; each
    # call returns once its frame is done (the frame ends with a stream
    # synchronisation that checks the frame's totals)
    per, total_ms = time_frames(r, scene, cam, args.steps, dev)
    ms_median = float(np.median(per))
    ms_mean = total_ms / args.steps
    # one untimed instrumented frame with the tuned options: rasterize work
    # units (pixel loop iterations) and the passes' binning sizes
    # (the loop count is the reference's, on 16-px tiles whatever the tuned
    # tile size), then one with the binning tallies at the tuned shape
    r.set_counting(True)
    r.render_scene(scene, cam, readback=False)
    cst = r.stats()
    r.set_counting("tallies")
    r.render_scene(scene, cam, readback=False)
    tst = r.stats()
    r.set_counting(False)
    iters, consumed = cst["pixel_iters"], cst["pairs_consumed"]
    note(f"instrumented frame: {iters} pixel iterations, {consumed} tile-list entries read; at "
         f"{tst['tile_size']}-px tiles: {tst['binned_splats']} binned splats, {tst['row_spans']} row spans, "
         f"{tst['group_spans']} group spans, {tst['list_entries']} list entries")
    # the same number of frames again with hipEvents at the stage boundaries
    # (each event leaves the GPU idle for a few microseconds)
    r.set_timing(True)
    stage_sum = {}
    for _ in range(args.steps):
        r.render_scene(scene, cam, readback=False)
        st = r.stats()
        for k, v in st["stage_ms"].items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    r.set_timing(False)
    stage_ms = {k: v / args.steps for k, v in stage_sum.items()}

    t1 = time.perf_counter()
    img = r.read_image()
    readback_ms = (time.perf_counter() - t1) * 1e3
    # into memory the caller already holds (no first-touch page faults)
    t1 = time.perf_counter()
    r.read_image(out=img)
    readback_reuse_ms = (time.perf_counter() - t1) * 1e3
    # a second resident copy of the scene: the upload alone (host arrays to
    # HBM through pinned staging, then the colour records built on the device)
    t1 = time.perf_counter()
    scene2 = Scene(dev, data)
    upload_ms = (time.perf_counter() - t1) * 1e3
    scene2.destroy()
    # the reference's own boundary: host arrays in, RGBA8 out (PCIe both ways);
    # the median of 3 frames (one frame of 5 has measured 20 ms slower)
    r.render(data, cam)
    host_per = []
    for _ in range(3):
        t1 = time.perf_counter()
        r.render(data, cam)
        host_per.append((time.perf_counter() - t1) * 1e3)
    host_ms = float(np.median(host_per))
    note(f"host-array render (upload + frame + readback): {host_ms:.1f} ms (median of 3)")
    st = r.stats()
    note(f"timed {args.steps} frames: median {ms_median:.3f} ms/frame (p10 {_pct(per, 10):.3f}, "
         f"p90 {_pct(per, 90):.3f}, mean {ms_mean:.3f})")

    rast_ms = stage_ms["rasterize"]
    rast_kernel = {(16, 1): "k_rasterize<false>", (16, 2): "k_rasterize_2x2<16>", (8, 1): "k_rasterize8",
                   (8, 2): "k_rasterize_2x2<8>", (32, 1): "k_rasterize32"}[(st["tile_size"], st["raster_group"])]
    achieved = iters * RAST_OPS_PER_ITER / (rast_ms * 1e-3) / 1e12
    traffic = _load_traffic()
    if traffic and traffic.get("kernel") != rast_kernel:
        traffic = None  # the committed counters are of another rasterizer
    line = {
        "metric": METRIC,
        "value": round(ms_median, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mean, 3),
        "p10_ms": round(_pct(per, 10), 3),
        "p90_ms": round(_pct(per, 90), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "large forward render: 6M Gaussians, SH bands 3, 3840x2160 pinhole, scene resident in HBM",
            "splats": data.count, "sh_bands": data.sh_bands, "width": cam.width, "height": cam.height,
            "projection": cam.projection, "visible": st["num_visible"], "pairs": st["num_pairs"],
            "tiles": st["num_tiles"], "tile_size": st["tile_size"], "raster_group": st["raster_group"],
            "reference_pairs_16px": cst["num_pairs"], "pairs_read_by_rasterize_16px": consumed,
        },
        "roofline": {
            "kernel": rast_kernel, "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            # one unit = the frame's rasterize stage: cst["num_chunks"] launches
            # (front-to-back passes); work, time and traffic all cover the
            # whole stage
            "traffic": traffic.get("bytes_per_frame") if traffic else None,
            "traffic_source": ("profiles/rasterize_traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of "
                               "this command's rasterize launches, scaled by a 1 GiB copy"
                               if traffic else None),
            "work": (f"{iters} pixel loop iterations x {RAST_OPS_PER_ITER} ops: the reference's loop (16-px "
                     "tiles), of which the kernel skips whole strips and exps it proves cannot change a pixel: "
                     "a work rate, not the FP32 instructions executed (those are in "
                     "profiles/rasterize_voxelize_valu_counters.md)"),
            "launches_per_frame": st["num_chunks"],
            "stage_ms": round(rast_ms, 4),
        },
        "autotune": tune,
        "stages": stage_table(stage_ms, tst, data.count, data.sh_bands),
        "readback_ms": round(readback_ms, 3),
        "readback_GB_s": round(img.nbytes / (readback_ms * 1e-3) / 1e9, 2),
        "readback_reused_buffer_ms": round(readback_reuse_ms, 3),
        "readback_reused_buffer_GB_s": round(img.nbytes / (readback_reuse_ms * 1e-3) / 1e9, 2),
        "scene_upload_ms": round(upload_ms, 3),
        "scene_upload_GB_s": round(data.nbytes() / (upload_ms * 1e-3) / 1e9, 2),
        "host_arrays_ms_per_frame": round(host_ms, 3),
        "image_nonblack_px": int(((img.reshape(-1, 4)[:, :3] > 5).any(1)).sum()),
    }
    if not args.no_equirect:
        line["equirect"] = equirect_leg(dev, scene, note)
    if not args.no_voxel:
        line["voxelize"] = voxelize_leg(dev, note, with_cpu=not args.no_cpu)
    if not args.no_batch:
        r.destroy()
        scene.destroy()
        line["batched"] = batched_views(dev, args.batch_steps, note, tune=not args.no_tune)
        r = scene = None
    if not args.no_cpu:
        note("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(data, cam)
    print(json.dumps(line))
    if r is not None:
        r.destroy()
        scene.destroy()
    dev.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
