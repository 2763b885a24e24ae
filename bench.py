#!/usr/bin/env python3
"""Incoherent closest-hit throughput on one MI355X.

Workload (the 1M-triangle heightfield configuration): 700x700 fBm heightfield + 2,000 boxes =
1,004,000 triangles (seed 0xC0FFEE02), SBVH built with the reference builder's semantics;
1920x1080 primary rays from (0.5, 0.6, -0.3) towards (0.5, 0.05, 0.6), 60 deg; one cosine-
weighted diffuse bounce per primary hit (tMin 1e-3). A step is one closest-hit launch over all
bounce rays (device resident). The launch shape is picked by the library's autotuner first.

Also reported (same GPU, same scene): the batched form -- 256 ray batches (sizes lognormal,
median 32k, 1k..2M; 40% primary / 40% diffuse / 20% shadow) as 256 plain launches, as one fused
persistent launch, and as its hipGraph replay, with identical outputs -- and the CPU baseline
(the reference's castRay compiled from its sources, on all host cores, bounded sample).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s incoherent closest-hit @1M tris; p50/p99 ray-batch latency"
HBM_PEAK_GBS = 8000.0


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


PROFILE = ROOT / "profiles" / "closest_persistent_bounce1M.md"
FUSED_PROFILE = ROOT / "profiles" / "trace_batches_fused256_1M.md"
# nominal L2 bandwidth: 8 XCDs x 16 channels x 128 B per clock at 2.4 GHz
L2_NOMINAL_GBS = 8 * 16 * 128 * 2.4


def shape_key(launch: dict) -> str:
    """grid,block,treelet,refill,helpers,raysPerFetch,ldsStack of a launch shape."""
    return ",".join(str(launch[k]) for k in ("gridBlocksClosest", "blockThreads", "topNodes",
                                              "refillMin", "helperWaves", "raysPerFetch",
                                              "stackLdsEntries"))


def profile_line(tag: str, path=PROFILE):
    """Text after `tag` on the first line of a committed rocprofv3 summary that holds it, or
    None."""
    try:
        for line in path.read_text().splitlines():
            if tag in line:
                return line.split(tag, 1)[1].strip()
    except OSError:
        pass
    return None


def profiled_traffic(key: str):
    """HBM bytes per launch of the closest-hit kernel on this workload under launch shape `key`,
    from the committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary (scaled by a plain copy's
    readings): lines `traffic_bytes_per_launch[<shape>]: <bytes>`. None if that shape was not
    profiled."""
    text = profile_line(f"traffic_bytes_per_launch[{key}]:")
    return int(text.split()[0].strip(" )`")) if text else None


def cpu_baseline(nodes, tris, rays, gpu_hits, budget_s: float = 24.0) -> dict:
    """The reference's own castRay (compiled from its sources into oracle/_ref) if present,
    else the oracle's C restatement, over a bounded, evenly strided sample of the rays:
    - value: the parity build (-O3 -ffp-contract=off) on all host cores pulling 1024-ray chunks,
      1 warm-up + median of 5;
    - with the reference build also: its own 10-thread fixed-band split
      (cpu_tracer_renderer/CPUTracerRenderer.cpp:77-87) and the as-shipped -ffast-math build
      (CMakeLists.txt:42) on all cores, median of 3 each, with the fast-math hits compared."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    use_ref = O.have_ref()
    cast = (lambda r: O.ref_cast(nodes, tris, r, threads)) if use_ref else \
        (lambda r: O.cast_closest(nodes, tris, r, threads))
    probe = rays[: min(len(rays), 20000)]
    t0 = time.perf_counter()
    cast(probe)
    rate = len(probe) / max(time.perf_counter() - t0, 1e-6)
    # the parity runs take about half the budget; the band split and fast-math runs the rest
    n = int(min(len(rays), max(len(probe), rate * budget_s / 12)))
    idx = np.linspace(0, len(rays) - 1, n).astype(np.int64)
    sample = np.ascontiguousarray(rays[idx])

    def timed(f, reps):
        times, out = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = f(sample)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), out

    cast(sample)   # warm-up
    t, hits = timed(cast, 5)
    same = hits.tobytes() == np.ascontiguousarray(gpu_hits[idx]).tobytes()
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"value": round(n / t / 1e6, 3), "unit": "Mrays/s", "cores": threads,
           "kind": "reference" if use_ref else "port",
           "sample": f"{n} of the {len(rays)} bounce rays (evenly strided), 1 warm-up + median of "
                     f"5 runs, "
                     f"{'castRay<144> compiled from the reference sources' if use_ref else 'oracle castRay restatement'}"
                     f" (-O3 -ffp-contract=off) on {threads} threads pulling 1024-ray chunks "
                     f"({cpu_model}); hits identical to GPU: {same}"}
    if use_ref:
        tb, hb = timed(lambda r: O.ref_cast_bands(nodes, tris, r, 10), 3)
        out["reference_10_thread_bands"] = {
            "value": round(n / tb / 1e6, 3), "cores": 10,
            "note": "the reference CPU renderer's split: 10 threads over fixed contiguous bands, "
                    f"median of 3; hits identical to GPU: "
                    f"{hb.tobytes() == np.ascontiguousarray(gpu_hits[idx]).tobytes()}"}
        if O.have_ref(fastmath=True):
            O.ref_cast(nodes, tris, sample[: min(n, 20000)], threads, fastmath=True)
            tf, hf = timed(lambda r: O.ref_cast(nodes, tris, r, threads, fastmath=True), 3)
            differ = int(np.count_nonzero(
                (hf["tri"] != hits["tri"]) | (hf["t"].view(np.uint32) != hits["t"].view(np.uint32))))
            out["reference_fastmath"] = {
                "value": round(n / tf / 1e6, 3), "cores": threads,
                "note": f"as-shipped -O3 -ffast-math build, median of 3; {differ} of {n} hits "
                        "differ from the parity build in triangle or t"}
    return out


def set_shape(ctx, p: dict) -> None:
    ctx.set_launch_params(grid_closest=p["gridBlocksClosest"], grid_anyhit=p["gridBlocksAnyhit"],
                          refill_min=p["refillMin"], block_threads=p["blockThreads"],
                          top_nodes=p["topNodes"], helper_waves=p["helperWaves"],
                          rays_per_fetch=p["raysPerFetch"], stack_lds_entries=p["stackLdsEntries"])


def batched_form(ctx, scene, pools, ev, reps: int = 5, plan_launch=None) -> dict:
    from phantasyenginetestbed_amd import phx, workloads as W
    bs = W.batch_set(pools)
    W.upload_batches(ctx, bs)
    stream = ctx.stream
    sizes = np.array([len(r) for r in bs.rays])
    log(f"batched form: {len(sizes)} batches, {bs.total} rays")

    # plain: one launch per batch, each timed with its own events
    evs = ev.create(len(sizes) + 1)
    W.launch_plain(ctx, scene, bs)
    ctx.synchronize()
    ref = W.download_outputs(bs)
    plain_best, plain_lat, plain_shape = None, None, None
    # The plain form gets the better of two shapes: the context's (tuned for one large cast) and
    # the library defaults (64 rays per fetch suit its small batches better); each shape's time is
    # the median of `reps` rounds after the warm-up above, with that round's per-batch latencies.
    tuned_shape = ctx.launch_params()
    default_shape = dict(tuned_shape, gridBlocksClosest=ctx.num_cus, gridBlocksAnyhit=ctx.num_cus,
                         refillMin=8, blockThreads=512, topNodes=0, helperWaves=0,
                         raysPerFetch=64, stackLdsEntries=16)
    for name, shape in (("tuned", tuned_shape), ("default", default_shape)):
        set_shape(ctx, shape)
        W.launch_plain(ctx, scene, bs)   # warm-up under this shape
        ctx.synchronize()
        rounds = []
        for _ in range(reps):
            ev.record(evs[0], stream)
            for i, (k, r, (rb, ob)) in enumerate(zip(bs.kinds, bs.rays, bs.device)):
                if k == "shadow":
                    ctx.cast_anyhit_device(scene, rb.ptr, len(r), ob.ptr, asynchronous=True)
                else:
                    ctx.cast_closest_device(scene, rb.ptr, len(r), ob.ptr, asynchronous=True)
                ev.record(evs[i + 1], stream)
            ctx.synchronize()
            lat = np.array([ev.elapsed_ms(evs[i], evs[i + 1]) for i in range(len(sizes))])
            rounds.append((ev.elapsed_ms(evs[0], evs[-1]), lat))
        rounds.sort(key=lambda x: x[0])
        tot, lat = rounds[len(rounds) // 2]
        if plain_best is None or tot < plain_best:
            plain_best, plain_lat, plain_shape = tot, lat, name
        same_plain = all(np.array_equal(a, b) for a, b in zip(ref, W.download_outputs(bs)))
        assert same_plain, f"plain outputs differ under the {name} shape"
    set_shape(ctx, tuned_shape)
    ev.destroy(evs)

    plan = W.plan_for(ctx, scene, bs, track=True)
    if plan_launch:
        grid, block, top, refill = plan_launch[:4]
        extra = list(plan_launch[4:]) + [0] * (3 - len(plan_launch[4:]))
        plan.set_launch_params(grid, block_threads=block, top_nodes=top, refill_min=refill,
                               helper_waves=extra[0], rays_per_fetch=extra[1],
                               stack_lds_entries=extra[2])
        tuned = {"gridBlocksClosest": grid, "blockThreads": block, "topNodes": top,
                 "refillMin": refill, "raysPerFetch": extra[1] or 64,
                 "stackLdsEntries": extra[2] or 16}
    else:
        tuned = plan.autotune()
    e2 = ev.create(2)

    def timed(graph: bool) -> float:
        # one warm-up launch, then the median of `reps` timed launches
        plan.launch(graph=graph)
        ms = []
        for _ in range(reps):
            ev.record(e2[0], stream)
            plan.launch(graph=graph, asynchronous=True)
            ev.record(e2[1], stream)
            ms.append(ev.elapsed_ms(e2[0], e2[1]))
        return float(np.median(ms))

    fused_ms = timed(False)
    fused_lat = plan.batch_times_ms()
    # algorithmic bytes of the fused launch: per batch, rays x (32 + out + 64 N_node + 48 N_tri)
    # with N from the traversal counters of an evenly strided 1/64 of the batch (closest-hit
    # counters for closest batches, any-hit traversal counters for shadow batches)
    algo = 0.0
    for k, r in zip(bs.kinds, bs.rays):
        sample = np.ascontiguousarray(r[::64])
        anyhit = k == "shadow"
        st = ctx.traversal_stats(scene, sample, phx.STATS_ANYHIT if anyhit else 0)
        per = 32 + (1 if anyhit else 16) + (64 * st["nodesVisited"] +
                                            48 * st["triangleIntersectionTests"]) / len(sample)
        algo += per * len(r)
    fused_gbs = algo / (fused_ms / 1e3) / 1e9
    same_fused = all(a.tobytes() == b.tobytes() for a, b in zip(ref, W.download_outputs(bs)))
    graph_ms = timed(True)
    same_graph = all(a.tobytes() == b.tobytes() for a, b in zip(ref, W.download_outputs(bs)))
    ev.destroy(e2)
    plan.release()
    W.free_batches(bs)
    return {
        "batches": int(len(sizes)), "rays": int(bs.total),
        "mix": {k: int(sum(len(r) for kk, r in zip(bs.kinds, bs.rays) if kk == k))
                for k in ("primary", "diffuse", "shadow")},
        "plain_ms": round(plain_best, 3), "plain_mrays": round(bs.total / plain_best / 1e3, 1),
        "plain_batch_p50_ms": round(float(np.percentile(plain_lat, 50)), 4),
        "plain_batch_p99_ms": round(float(np.percentile(plain_lat, 99)), 4),
        "fused_ms": round(fused_ms, 3), "fused_mrays": round(bs.total / fused_ms / 1e3, 1),
        "fused_graph_ms": round(graph_ms, 3),
        "fused_graph_mrays": round(bs.total / graph_ms / 1e3, 1),
        "fused_batch_completion_p50_ms": round(float(np.percentile(fused_lat, 50)), 4),
        "fused_batch_completion_p99_ms": round(float(np.percentile(fused_lat, 99)), 4),
        "fused_vs_plain": round(plain_best / fused_ms, 2),
        "fused_roofline": {"bound": "issue", "achieved": round(fused_gbs, 1), "unit": "GB/s",
                           "hbm_peak": HBM_PEAK_GBS,
                           "frac_of_hbm_peak": round(fused_gbs / HBM_PEAK_GBS, 4),
                           "l2_nominal_peak": round(L2_NOMINAL_GBS, 1),
                           "frac_of_l2_nominal": round(fused_gbs / L2_NOMINAL_GBS, 4),
                           "algorithmic_bytes": int(algo),
                           "limiter": profile_line("limiter:", FUSED_PROFILE),
                           "note": "lane-visit bytes (per-kind node/triangle counts on a strided "
                                   "1/64 of every batch); coherent primary batches share node "
                                   "loads in L1/L2, so they can exceed what HBM delivers"},
        "plain_shape": plain_shape,
        "identical_outputs": bool(same_fused and same_graph),
        "fused_launch": {k: tuned[k] for k in ("gridBlocksClosest", "blockThreads", "topNodes",
                                               "refillMin", "raysPerFetch", "stackLdsEntries")},
    }


def frame_form(ctx, scene, ev, lights: int, frames: int = 16) -> dict:
    """The wavefront frame at 1920x1080 with TAA over 16 Halton-jittered frames, replayed from the
    captured hipGraph (steady frames), beside the same frames launched directly; identical
    outputs."""
    from phantasyenginetestbed_amd import phx, workloads as W
    w, h = 1920, 1080
    cam = W.camera(w, h)
    outs = {}

    def run(extra):
        r = ctx.renderer(w, h)
        out = ctx.alloc(w * h * 16)
        flags = phx.FRAME_TAA | extra
        for f in range(2):  # warm-up (allocates the per-light buffers, captures the graph)
            r.render(scene, cam, f, 3, flags | phx.FRAME_TAA_RESET, out_ptr=out.ptr)
        evs = ev.create(frames + 1)
        ev.record(evs[0], ctx.stream)
        for f in range(frames):
            r.render(scene, cam, f, 3, flags | (phx.FRAME_TAA_RESET if f == 0 else 0),
                     out_ptr=out.ptr, asynchronous=True)
            ev.record(evs[f + 1], ctx.stream)
        ctx.synchronize()
        per = np.array([ev.elapsed_ms(evs[i], evs[i + 1]) for i in range(frames)])
        ev.destroy(evs)
        hist = r.download(phx.BUFFER_HISTORY, np.float32).reshape(h, w, 4)
        outs[extra] = out.download(np.float32, w * h * 4)
        r.release()
        out.free()
        return per, hist

    direct, _ = run(0)
    per, hist = run(phx.FRAME_GRAPH)
    n, m = w * h, (w // 2) * (h // 2)
    rays = n + n * lights + m + m * lights
    ms = float(per.mean())
    return {"resolution": [w, h], "lights": lights, "materials": 64, "frames": frames,
            "ms_per_frame": round(ms, 3), "p50_ms": round(float(np.percentile(per, 50)), 3),
            "p99_ms": round(float(np.percentile(per, 99)), 3),
            "rays_per_frame": int(rays), "mrays": round(rays / ms / 1e3, 1),
            "replay": "hipGraph", "direct_ms_per_frame": round(float(direct.mean()), 3),
            "identical_outputs": bool(outs[0].tobytes() == outs[phx.FRAME_GRAPH].tobytes()),
            "taa_samples_after": float(hist[..., 3].min())}


def overlapped_form(ctx, scene, bvh, rays, n: int, want: np.ndarray, launches: int = 20) -> dict:
    """The headline batches with consecutive batches overlapping instead of one launch after the
    other on one stream: (a) the same persistent launches alternating between two contexts (two
    streams, each with its own ray counters), so one launch's tail runs beside the next launch;
    (b) the same batches as device-resident requests to the request server (3 slot streams,
    fused multi-request launches), all submitted at once, with each request's latency. The median
    of 3 runs each; every output equals the one-stream hits."""
    from phantasyenginetestbed_amd import phx
    lp = ctx.launch_params()
    other = phx.Context(0)
    oscene = other.upload_scene(bvh.nodes, bvh.tris, bvh.datas)
    other.set_launch_params(grid_closest=lp["gridBlocksClosest"], grid_anyhit=lp["gridBlocksAnyhit"],
                            refill_min=lp["refillMin"], block_threads=lp["blockThreads"],
                            top_nodes=lp["topNodes"], helper_waves=lp["helperWaves"],
                            rays_per_fetch=lp["raysPerFetch"],
                            stack_lds_entries=lp["stackLdsEntries"])
    outs = [ctx.alloc(n * 16) for _ in range(launches)]
    try:
        pair = [(ctx, scene, outs[0]), (other, oscene, outs[1])]
        for c, sc, o in pair:  # warm-up
            c.cast_closest_device(sc, rays.ptr, n, o.ptr)
        runs = []
        for _ in range(3):
            ctx.synchronize()
            other.synchronize()
            t0 = time.perf_counter()
            for k in range(launches):
                c, sc, o = pair[k % 2]
                c.cast_closest_device(sc, rays.ptr, n, o.ptr, asynchronous=True)
            ctx.synchronize()
            other.synchronize()
            runs.append(time.perf_counter() - t0)
        two = launches * n / float(np.median(runs)) / 1e6
        same_two = all(o.download(phx.HIT, n).tobytes() == want.tobytes() for _, _, o in pair)
        oscene.release()
        other.destroy()
        other = None
        srv = phx.Server(ctx, 4 << 30, slots=3)
        try:
            srv.add_scene(scene)
            reqs = [phx.Server.request(scene, rays.ptr, o.ptr, n, phx.BATCH_CLOSEST, device=True)
                    for o in outs]
            srv.replay(reqs, np.zeros(launches))  # warm-up
            rounds = [srv.replay(reqs, np.zeros(launches)) for _ in range(3)]
        finally:
            srv.destroy()
        lat, mk = sorted(rounds, key=lambda r: r[1])[1]
        same_srv = all(o.download(phx.HIT, n).tobytes() == want.tobytes() for o in outs)
    finally:
        if other is not None:
            oscene.release()
            other.destroy()
        for o in outs:
            o.free()
    return {"batches": launches, "rays_per_batch": int(n),
            "two_streams": {"mrays": round(two, 1), "identical_hits": bool(same_two),
                            "note": "persistent launches alternating between two contexts"},
            "served_device_resident": {"mrays": round(launches * n / mk / 1e3, 1),
                                       "makespan_ms": round(mk, 3),
                                       "batch_p50_ms": round(float(np.percentile(lat, 50)), 3),
                                       "batch_p99_ms": round(float(np.percentile(lat, 99)), 3),
                                       "slots": 3, "identical_hits": bool(same_srv),
                                       "note": "all batches submitted at once; latency is "
                                               "submit to completion, queueing included"},
            "note": "not the headline value: the same batches with their launches overlapping, "
                    "so one launch's completion tail runs beside the next batch's rays"}


def instances_form(ctx, reps: int = 5) -> dict:
    """Dynamic meshes: 1M rays against 300 instances of three meshes under a top tree
    (phx_cast_closest_instances / phx_cast_anyhit_instances, device buffers), median of `reps`
    casts each after a warm-up."""
    from phantasyenginetestbed_amd import phx, workloads as W
    wl = W.instance_workload(ctx)
    n = len(wl.rays)
    iset = phx.InstanceSet(ctx, wl.instances)
    rb = ctx.alloc(n * 32).upload(wl.rays)
    hb, ib, lb = ctx.alloc(n * 16), ctx.alloc(n * 4), ctx.alloc(n)
    try:
        def closest():
            iset.cast_closest_device(rb.ptr, n, hb.ptr, ib.ptr)

        def anyhit():
            phx.check(phx.lib().phx_cast_anyhit_instances(ctx.handle, None, iset.handle, rb.ptr,
                                                          n, lb.ptr, phx.MEM_DEVICE))
        out = {"instances": len(wl.instances), "mesh_tris": [300, 4000, 20000], "rays": n}
        for name, fn in (("closest", closest), ("anyhit", anyhit)):
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            ms = float(np.median(ts)) * 1e3
            out[name] = {"ms": round(ms, 3), "mrays": round(n / ms / 1e3, 1)}
        hits = hb.download(phx.HIT, n)
        out["closest"]["hit_fraction"] = round(float((hits["tri"] != phx.MISS).mean()), 4)
        out["anyhit"]["occluded_fraction"] = round(float((lb.download(np.uint8, n) == 0).mean()), 4)
    finally:
        iset.release()
        for b in (rb, hb, ib, lb):
            b.free()
        for m in wl.meshes:
            m.release()
    return out

def served_form(ctx, n_requests: int = 1500, budget_gb: float = 64.0, slots: int = 3) -> dict:
    """Poisson-arriving requests against 4 resident scenes through the request server, at
    saturation (all requests at once; after a warm-up, the median of 5 runs by makespan) and at
    25/50/90% of the saturated request rate (3 runs each with their own arrival draws; medians of
    the runs' rate, p50 and p99), beside the plain per-batch server on the same requests (median of
    5 saturated runs; identical outputs). The same requests with device-resident rays and outputs
    (no PCIe copies) show the engine's own serving rate."""
    from phantasyenginetestbed_amd import phx, workloads as W
    t0 = time.perf_counter()
    scenes = [W.served_scene(ctx, name) for name in ("100k", "1M", "4M", "8M")]
    log(f"served: 4 scenes ({', '.join(f'{s.name}={s.triangles}' for s in scenes)}) "
        f"in {time.perf_counter() - t0:.1f}s")
    rs = W.request_set(len(scenes), n_requests)
    W.allocate_outputs(rs)
    reqs = W.server_requests(scenes, rs)
    total = rs.total

    def run(server, arrivals, which=None):
        lat, mk = server.replay(which or reqs, arrivals)
        return {"mrays": round(total / mk / 1e3, 1), "makespan_ms": round(mk, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3)}

    srv = phx.Server(ctx, int(budget_gb * (1 << 30)), slots=slots)
    for s in scenes:
        srv.add_scene(s.scene)
    zeros = np.zeros(n_requests)

    def median_run(server, reps=5, which=None):
        runs = sorted((run(server, zeros, which) for _ in range(reps)),
                      key=lambda r: r["makespan_ms"])
        return runs[len(runs) // 2]

    run(srv, zeros)  # warm-up
    sat = median_run(srv)
    served_out = rs.out.array.copy()
    rate = n_requests / sat["makespan_ms"]  # requests per ms at saturation
    rates = {}
    for frac in (0.25, 0.5, 0.9):
        runs = [run(srv, W.poisson_arrivals(n_requests, frac * rate, seed=int(frac * 100) + 7 * k))
                for k in range(3)]
        med = {key: float(np.median([r[key] for r in runs]))
               for key in ("mrays", "makespan_ms", "p50_ms", "p99_ms")}
        rates[f"{int(frac * 100)}%"] = {"offered_req_per_s": round(frac * rate * 1e3, 1),
                                        **{k: round(v, 3) for k, v in med.items()},
                                        "runs_p50_ms": [r["p50_ms"] for r in runs],
                                        "runs_p99_ms": [r["p99_ms"] for r in runs]}
    st = srv.stats()
    # device-resident requests: the same mix with rays and outputs in HBM
    dreqs, keep, dout = W.server_requests_device(ctx, scenes, rs)
    run(srv, zeros, dreqs)  # warm-up
    dsat = median_run(srv, which=dreqs)
    drate = n_requests / dsat["makespan_ms"]
    dpois = run(srv, W.poisson_arrivals(n_requests, 0.9 * drate, seed=90), dreqs)
    same_device = bool(np.array_equal(dout.download(np.uint8, len(served_out)), served_out))
    for b in keep:
        b.free()
    srv.destroy()

    plain = phx.Server(ctx, int(budget_gb * (1 << 30)), slots=1, plain=True)
    for s in scenes:
        plain.add_scene(s.scene)
    run(plain, zeros)
    psat = median_run(plain)
    same = bool(np.array_equal(served_out, rs.out.array))
    plain.destroy()
    rs.out.free()
    for s in scenes:
        W.free_served_scene(s)
    return {"scenes": {s.name: s.triangles for s in scenes}, "requests": n_requests,
            "rays": total, "budget_gb": budget_gb, "slots": slots,
            "served_saturated": sat, "served_at_rates": rates, "plain_saturated": psat,
            "served_vs_plain": round(sat["mrays"] / psat["mrays"], 2),
            "device_resident": {"saturated": dsat, "at_90%": {
                "offered_req_per_s": round(0.9 * drate * 1e3, 1), **dpois},
                "identical_outputs": same_device,
                "note": "rays and outputs in HBM: the engine's serving rate without PCIe copies"},
            "identical_outputs": same, "launches": int(st["launches"]),
            "peak_request_bytes": int(st["peakBytesInUse"]),
            "note": "host-resident (pinned) request buffers: times include PCIe copies"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=700)
    ap.add_argument("--boxes", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-overlap", action="store_true")
    ap.add_argument("--no-instances", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default=None,
                    metavar="GRID,BLOCK,TOP,REFILL[,HELPERS[,RAYS_PER_FETCH[,LDS_STACK]]]",
                    help="fixed persistent launch shape instead of the autotuner's (profiling)")
    ap.add_argument("--plan-launch", default=None,
                    metavar="GRID,BLOCK,TOP,REFILL[,HELPERS[,RAYS_PER_FETCH[,LDS_STACK]]]",
                    help="fixed launch shape of the fused 256-batch plan (profiling)")
    ap.add_argument("--kernel", choices=["persistent", "one-ray-per-lane", "packet"], default=None,
                    help="force a launch model (default: the autotuner's choice)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first (a plain copy kernel "
                         "of known size, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the rays of a batch "
              "shard across GPUs with no exchange step (each GPU holds the scene); this benchmark "
              "runs on a single MI355X", file=sys.stderr)
        return 2

    from phantasyenginetestbed_amd import phx, workloads as W

    log(f"scene {args.grid}x{args.grid} heightfield + {args.boxes} boxes")
    t0 = time.perf_counter()
    tris, datas, bvh = W.heightfield_scene(args.grid, args.boxes, materials=64)
    log(f"SBVH: {len(tris)} tris -> {len(bvh.nodes)} nodes, {len(bvh.tris)} refs, "
        f"depth {bvh.depth()}, {time.perf_counter() - t0:.2f}s")

    ctx = phx.Context(0)
    lights = phx.lights_random(8, 0xC0FFEE03)
    scene = ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas, phx.materials_random(64, 0xC0FFEE03),
                             lights)
    pools = W.ray_pools(ctx, scene)
    bounce = pools.bounce
    n = len(bounce)
    rays = ctx.alloc(n * 32).upload(bounce)
    out = ctx.alloc(n * 16)
    log(f"{n} bounce rays from {len(pools.primary)} primary rays")

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

    launch = {"kernel": "persistent", **ctx.launch_params()}
    if args.launch:
        vals = [int(x) for x in args.launch.split(",")]
        grid, block, top, refill = vals[:4]
        extra = vals[4:] + [0] * (3 - len(vals[4:]))
        ctx.set_launch_params(grid_closest=grid, grid_anyhit=grid, refill_min=refill,
                              block_threads=block, top_nodes=top, helper_waves=extra[0],
                              rays_per_fetch=extra[1], stack_lds_entries=extra[2])
        launch = {"kernel": "persistent", **ctx.launch_params()}
        log(f"fixed launch: {launch}")
    elif not args.no_autotune:
        launch = ctx.autotune(scene, rays.ptr, n)
        log(f"autotuned launch: {launch}")
    kernel = args.kernel or launch["kernel"]
    mode = {"one-ray-per-lane": phx.KERNEL_ONE_RAY_PER_LANE,
            "packet": phx.KERNEL_PACKET}.get(kernel, 0)

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
    tail = None
    if kernel == "persistent":
        # one more launch of the same shape that also records when each ray finished
        done = ctx.cast_closest_timeline(scene, rays.ptr, n, out.ptr)
        same = out.download(phx.HIT, n).tobytes() == gpu_hits.tobytes()
        q = np.sort(done)
        tail = {"rays_done_ms": {f"{f:g}%": round(float(q[min(n - 1, int(np.ceil(f / 100 * n)) - 1)]), 4)
                                 for f in (50, 90, 99, 99.9)},
                "last_ms": round(float(q[-1]), 4),
                "identical_hits": bool(same),
                "note": "one extra launch (per-ray device clock at each result write), "
                        "phx_cast_closest_timeline"}
        log(f"tail: {tail}")
    value = n / (ms_per_step / 1e3) / 1e6
    achieved = bytes_per_ray * n / (ms_per_step / 1e3) / 1e9
    key = shape_key(launch)
    traffic = profiled_traffic(key) if kernel == "persistent" else None
    if traffic is not None:
        traffic_source = (f"rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of this launch shape "
                          f"[{key}] (grid,block,treelet,refill,helpers,raysPerFetch,ldsStack; "
                          "profiles/closest_persistent_bounce1M.md)")
    else:
        traffic_source = f"no rocprofv3 profile of this launch shape [{key}]"

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "diffuse-bounce closest-hit, 1,004,000-tri heightfield+boxes "
                               "SBVH, rays from 1920x1080 primary hits, one launch per step",
                   "scene_tris": int(len(tris)), "sbvh_nodes": int(len(bvh.nodes)),
                   "sbvh_refs": int(len(bvh.tris)), "rays_per_step": int(n),
                   "traversal": "reference-order (bit-identical to castRay)",
                   "kernel": kernel,
                   "launch": {k: launch[k] for k in ("gridBlocksClosest", "blockThreads",
                                                     "topNodes", "refillMin", "helperWaves",
                                                     "raysPerFetch", "stackLdsEntries")
                              if k in launch}},
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "tail": tail,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_source,
                     "limiter": profile_line("limiter:"),
                     "algorithmic_bytes_per_ray": round(bytes_per_ray, 1)},
    }
    if not args.no_overlap and kernel == "persistent":
        try:  # a side measurement: its failure must not cost the headline line
            result["overlapped"] = overlapped_form(ctx, scene, bvh, rays, n, gpu_hits)
        except phx.PhxError as e:
            result["overlapped"] = {"error": str(e)}
        log(f"overlapped: {result['overlapped']}")
    if not args.no_instances:
        try:  # a side measurement: its failure must not cost the headline line
            result["instances"] = instances_form(ctx)
        except phx.PhxError as e:
            result["instances"] = {"error": str(e)}
        log(f"instances: {result['instances']}")
    if not args.no_batched:
        plan_launch = tuple(int(x) for x in args.plan_launch.split(",")) if args.plan_launch \
            else None
        result["batched"] = batched_form(ctx, scene, pools, ev, plan_launch=plan_launch)
        log(f"batched: {result['batched']}")
    if not args.no_frame:
        result["frame"] = frame_form(ctx, scene, ev, len(lights))
        log(f"frame: {result['frame']}")
    if not args.no_served:
        result["served"] = served_form(ctx)
        log(f"served: {result['served']}")
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
