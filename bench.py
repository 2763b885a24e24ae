#!/usr/bin/env python3
"""Incoherent closest-hit throughput on one MI355X.

Workload: the 1,004,000-triangle heightfield scene (700x700 fBm grid + 2,000 boxes, seed
0xC0FFEE02), SBVH built with the reference's builder semantics. 1920x1080 primary rays from
(0.5, 0.6, -0.3) towards (0.5, 0.05, 0.6), 60 deg; every primary hit spawns one cosine-weighted
diffuse bounce ray (tMin 1e-3). A step is one closest-hit pass over all bounce rays (device
resident). Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s incoherent closest-hit @1M tris; p50/p99 ray-batch latency"
HBM_PEAK_GBS = 8000.0
CAMERA = ((0.5, 0.6, -0.3), (0.5, 0.05, 0.6), (0.0, 1.0, 0.0))
SCENE_SEED = 0xC0FFEE02
BOUNCE_SEED = 0xC0FFEE02


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on the context's stream (libamdhip64 through ctypes)."""

    def __init__(self):
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        self.L = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self, n: int) -> list:
        evs = []
        for _ in range(n):
            e = C.c_void_p()
            assert self.L.hipEventCreate(C.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, e, stream) -> None:
        assert self.L.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b) -> float:
        assert self.L.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)

    def destroy(self, evs) -> None:
        for e in evs:
            self.L.hipEventDestroy(e)


def cpu_baseline(nodes, tris, rays, gpu_hits, budget_s: float = 12.0) -> dict:
    """The reference's own castRay (compiled from its sources into oracle/_ref) if present,
    else the oracle's C restatement, on all host cores over a bounded sample of the rays."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    use_ref = O.have_ref()
    cast = (lambda r: O.ref_cast(nodes, tris, r, threads)) if use_ref else \
        (lambda r: O.cast_closest(nodes, tris, r, threads))
    probe = rays[: min(len(rays), 20000)]
    t0 = time.perf_counter()
    cast(probe)
    rate = len(probe) / max(time.perf_counter() - t0, 1e-6)
    n = int(min(len(rays), max(len(probe), rate * budget_s / 2)))
    idx = np.linspace(0, len(rays) - 1, n).astype(np.int64)
    sample = np.ascontiguousarray(rays[idx])
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        hits = cast(sample)
        times.append(time.perf_counter() - t0)
    t = min(times)
    same = hits.tobytes() == np.ascontiguousarray(gpu_hits[idx]).tobytes()
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(n / t / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"{n} of the {len(rays)} bounce rays (evenly strided), best of 2 runs, "
                      f"{'castRay<144> compiled from the reference sources' if use_ref else 'oracle castRay restatement'}"
                      f" on {threads} threads ({cpu_model}); hits identical to GPU: {same}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=700)
    ap.add_argument("--boxes", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["persistent", "one-ray-per-lane"], default="persistent")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first (a plain copy kernel "
                         "of known size, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (rays shard with "
              "no exchange, see DESIGN.md; this build runs on a single MI355X)", file=sys.stderr)
        return 2

    from phantasyenginetestbed_amd import phx

    log(f"scene {args.grid}x{args.grid} heightfield + {args.boxes} boxes")
    tris, datas = phx.heightfield_boxes(args.grid, args.boxes, SCENE_SEED)
    t0 = time.perf_counter()
    bvh = phx.BVH(tris, datas)
    log(f"SBVH: {len(tris)} tris -> {len(bvh.nodes)} nodes, {len(bvh.tris)} refs, "
        f"depth {bvh.depth()}, {time.perf_counter() - t0:.2f}s")

    ctx = phx.Context(0)
    scene = ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas)
    W, H = 1920, 1080
    n_px = W * H
    cam = phx.camera_look_at(*CAMERA, math.radians(60.0), W, H)
    prim = ctx.alloc(n_px * 32)
    phit = ctx.alloc(n_px * 16)
    bnc = ctx.alloc(n_px * 32)
    ctx.gen_primary_rays_device(cam, W, H, prim.ptr)
    ctx.cast_closest_device(scene, prim.ptr, n_px, phit.ptr)
    ctx.gen_bounce_rays_device(scene, prim.ptr, phit.ptr, n_px, BOUNCE_SEED, 0, bnc.ptr)
    hits = phit.download(phx.HIT, n_px)
    bounce_all = bnc.download(phx.RAY, n_px)
    bounce = np.ascontiguousarray(bounce_all[hits["tri"] != phx.MISS])
    n = len(bounce)
    rays = ctx.alloc(n * 32).upload(bounce)
    out = ctx.alloc(n * 16)
    for b in (prim, phit, bnc):
        b.free()
    log(f"{n} bounce rays from {n_px} primary rays")

    counters = ctx.traversal_stats_device(scene, rays.ptr, n)
    nodes_per_ray = counters["nodesVisited"] / n
    tris_per_ray = counters["triangleIntersectionTests"] / n
    bytes_per_ray = 32 + 16 + 64 * nodes_per_ray + 48 * tris_per_ray
    log(f"per ray: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} triangle tests, "
        f"{bytes_per_ray:.0f} algorithmic bytes")

    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_on_device(dst.ptr, src.ptr, nb)
        src.free()
        dst.free()
        log(f"calibration copy of {args.calibrate_copy} MiB done")

    mode = phx.KERNEL_ONE_RAY_PER_LANE if args.kernel == "one-ray-per-lane" else 0
    ev = HipEvents()
    stream = ctx.stream
    for _ in range(args.warmup):
        ctx.cast_closest_device(scene, rays.ptr, n, out.ptr, mode=mode, asynchronous=True)
    ctx.synchronize()
    evs = ev.create(args.steps + 1)
    t0 = time.perf_counter()
    ev.record(evs[0], stream)
    for k in range(args.steps):
        ctx.cast_closest_device(scene, rays.ptr, n, out.ptr, mode=mode, asynchronous=True)
        ev.record(evs[k + 1], stream)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [ev.elapsed_ms(evs[k], evs[k + 1]) for k in range(args.steps)]
    ev.destroy(evs)
    total_ms = sum(step_ms)
    ms_per_step = total_ms / args.steps
    log(f"{args.steps} steps: events {total_ms:.2f} ms, wall {wall * 1e3:.2f} ms")

    gpu_hits = out.download(phx.HIT, n)
    value = n / (ms_per_step / 1e3) / 1e6
    achieved = bytes_per_ray * n / (ms_per_step / 1e3) / 1e9

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "diffuse-bounce closest-hit, 1,004,000-tri heightfield+boxes "
                               "SBVH, rays from 1920x1080 primary hits",
                   "scene_tris": int(len(tris)), "sbvh_nodes": int(len(bvh.nodes)),
                   "sbvh_refs": int(len(bvh.tris)), "rays_per_step": int(n),
                   "traversal": "reference-order (bit-identical to castRay)",
                   "kernel": args.kernel},
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "algorithmic_bytes_per_ray": round(bytes_per_ray, 1)},
    }
    if not args.no_cpu:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(bvh.nodes, bvh.tris, bounce, gpu_hits)
    print(json.dumps(result), flush=True)
    out.free()
    rays.free()
    scene.release()
    ctx.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
