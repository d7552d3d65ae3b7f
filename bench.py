"""Benchmark of the forward splat renderer on one MI355X.

Workload: 6,000,000 Gaussians (positions ~N(0, 0.75^2) clipped to [-3, 3],
SH bands 3), 3840x2160 pinhole from (0,0,4), scene resident in HBM. A step is
one full frame: f64 cull + depth keys, depth sort, project, scan, emit, pair
sort, tile ranges, rasterize + finalize to RGBA8 in device memory. The RGBA8
readback is timed separately and is not part of `value`.

Prints one JSON line (metric/value in ms per frame, lower is better) with:
  roofline      dominant kernel (rasterize) against the MI355X FP32 peak,
                duration from hipEvents on the renderer's stream
  cpu_baseline  the CPU oracle on the same scene: every stage on all 6M
                splats, rasterize on every ROW_STEP-th tile row scaled up
  batched       512 views (8 scenes x 64 orbit cameras) as one batched call
                against the same views rendered one by one
  voxelize      the 1M-Gaussian voxelizer workload (device time and wall
                time with the host transfers)
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
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (published)
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E (published)
# Algorithmic work per rasterize loop iteration: the reference's per-pixel
# body (rasterize.ts:62-79) is 31 f32 operations counting each add, mul,
# compare, min/max and the exp as one.
RAST_OPS_PER_ITER = 31
ROW_STEP = 16                # CPU baseline: rasterize every 16th tile row


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


def batched_views(dev, steps: int, note, tune: bool = True):
    """8 resident scenes (50k .. 4M Gaussians, render distributions) x 64
    orbit cameras at 640x360 / 1280x720 / 1920x1080: 512 views. Plain
    one-by-one renders of every view against one batched call (graph
    replay); a sample of views is checked bit-exact between the two."""
    from droplet_amd import synthetic as S
    from droplet_amd.batch import BatchRenderer, View
    from droplet_amd.renderer import Renderer, Scene

    datas, views = S.batch_workload()
    scenes = [Scene(dev, d) for d in datas]
    del datas
    r = Renderer(dev)
    for v in views[::37]:
        r.render_scene(scenes[v.scene], v.camera, readback=False)
    t0 = time.perf_counter()
    for v in views:
        r.render_scene(scenes[v.scene], v.camera, readback=False)
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
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render(scenes, views)
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


def voxelize_leg(dev, note):
    """The voxelizer workload: 1M Gaussians uniform in a 10 x 3 x 10 m box,
    voxelResolution 0.02, opacityCutoff 0.1 (host arrays in, block list out,
    as voxelize() takes and returns them)."""
    from droplet_amd import synthetic as S
    from droplet_amd.voxel import VoxelizeOptions, voxelize

    data, o = S.voxel_workload()
    opts = VoxelizeOptions(voxel_resolution=o["voxel_resolution"], opacity_cutoff=o["opacity_cutoff"])
    voxelize(dev, data, opts)  # warm
    walls, devs = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        res = voxelize(dev, data, opts)
        walls.append((time.perf_counter() - t0) * 1e3)
        devs.append(res.device_ms)
    note(f"voxelize: {min(devs):.3f} ms device, {min(walls):.3f} ms wall, {res.buffer.count} blocks")
    return {"workload": "1M Gaussians in a 10x3x10 m box, voxelResolution 0.02, opacityCutoff 0.1",
            "device_ms": round(min(devs), 3), "wall_ms_incl_pcie": round(min(walls), 3),
            "blocks": int(res.buffer.count), "solid_voxels": int(res.buffer.popcount()),
            "gaussians_per_s": round(data.count / (min(devs) * 1e-3), 1),
            "collision": collision_leg(dev, data, opts, note)}


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--splats", type=int, default=6_000_000)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-views leg")
    ap.add_argument("--no-tune", action="store_true", help="time the default options (no autotune)")
    ap.add_argument("--no-voxel", action="store_true", help="skip the voxelizer leg")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB device-to-device copy (a kernel of known traffic) so that a "
                         "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE run can calibrate the counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the renderer runs on a single "
              "device and multi-GPU sharding is not part of this build.", file=sys.stderr)
        return 2

    from droplet_amd import synthetic as S
    from droplet_amd.renderer import Renderer, Scene, create_device

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
    scene = Scene(dev, data)
    r = Renderer(dev)
    note(f"scene resident on {dev.name}")

    # one untimed instrumented frame: rasterize work units (pixel loop iterations)
    r.set_counting(True)
    r.render_scene(scene, cam, readback=False)
    cst = r.stats()
    iters = cst["pixel_iters"]
    consumed = cst["pairs_consumed"]
    r.set_counting(False)
    note(f"instrumented frame: {iters} pixel iterations, {consumed} tile-list entries read")

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
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render_scene(scene, cam, readback=False)
    wall = time.perf_counter() - t0
    ms_per_frame = wall * 1e3 / args.steps
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
    stage_ms = {k: round(v / args.steps, 4) for k, v in stage_sum.items()}

    t1 = time.perf_counter()
    img = r.read_image()
    readback_ms = (time.perf_counter() - t1) * 1e3
    # the reference's own boundary: host arrays in, RGBA8 out (PCIe both ways)
    r.render(data, cam)
    t1 = time.perf_counter()
    r.render(data, cam)
    host_ms = (time.perf_counter() - t1) * 1e3
    note(f"host-array render (upload + frame + readback): {host_ms:.1f} ms")
    st = r.stats()
    note(f"timed {args.steps} frames: {ms_per_frame:.3f} ms/frame")

    rast_ms = stage_ms["rasterize"]
    achieved = iters * RAST_OPS_PER_ITER / (rast_ms * 1e-3) / 1e12
    traffic = _load_traffic()
    line = {
        "metric": METRIC,
        "value": round(ms_per_frame, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "large forward render: 6M Gaussians, SH bands 3, 3840x2160 pinhole, scene resident in HBM",
            "splats": data.count, "sh_bands": data.sh_bands, "width": cam.width, "height": cam.height,
            "projection": cam.projection, "visible": st["num_visible"], "pairs": st["num_pairs"],
            "tiles": st["num_tiles"], "pairs_read_by_rasterize": consumed,
        },
        "roofline": {
            "kernel": "k_rasterize", "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            # one unit = the frame's rasterize stage: st["num_chunks"] launches
            # (front-to-back passes); work, time and traffic all cover the
            # whole stage
            "traffic": traffic.get("bytes_per_frame") if traffic else None,
            "work": f"{iters} pixel loop iterations x {RAST_OPS_PER_ITER} ops",
            "launches_per_frame": st["num_chunks"],
            "stage_ms": rast_ms,
        },
        "autotune": tune,
        "stage_ms": stage_ms,
        "readback_ms": round(readback_ms, 3),
        "host_arrays_ms_per_frame": round(host_ms, 3),
        "image_nonblack_px": int(((img.reshape(-1, 4)[:, :3] > 5).any(1)).sum()),
    }
    if not args.no_voxel:
        line["voxelize"] = voxelize_leg(dev, note)
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
