"""Synthetic workloads of the benchmark configurations, built on the C ABI.

  heightfield scene   700x700 fBm heightfield + 2,000 boxes (1,004,000 triangles), seed 0xC0FFEE02
                      (the 512x512 grid gives the 548,288-triangle scene)
  camera              from (0.5, 0.6, -0.3) towards (0.5, 0.05, 0.6), 60 deg vertical FOV,
                      up orthogonalised as ViewFrustum::set does
  ray pools           1920x1080 primary rays; one cosine-weighted diffuse bounce per primary hit
                      (tMin 1e-3); one shadow ray per primary hit to the point light (0.3, 2, 0.4),
                      tMin 1e-3, tMax = distance - 1e-3
  batched form        256 batches, sizes lognormal(ln 32768, 1.5) clipped to [1024, 2097152],
                      kinds 40% primary / 40% diffuse / 20% shadow (seed 0xC0FFEE04), each a
                      contiguous slice (wrapping) of its pool
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import phx

SCENE_SEED = 0xC0FFEE02
BOUNCE_SEED = 0xC0FFEE02
BATCH_SEED = 0xC0FFEE04
CAMERA = ((0.5, 0.6, -0.3), (0.5, 0.05, 0.6), (0.0, 1.0, 0.0))
LIGHT = (0.3, 2.0, 0.4)
FOV_DEG = 60.0


def heightfield_scene(grid: int = 700, boxes: int = 2000, materials: int = 0):
    tris, datas = phx.heightfield_boxes(grid, boxes, SCENE_SEED, num_materials=materials)
    return tris, datas, phx.BVH(tris, datas)


def camera(width: int = 1920, height: int = 1080) -> phx.CameraDef:
    return phx.camera_look_at(*CAMERA, math.radians(FOV_DEG), width, height)


def shadow_rays(rays: np.ndarray, hits: np.ndarray, light=LIGHT) -> np.ndarray:
    """Shadow ray from every hit to a point light: tMin 1e-3, tMax = distance - 1e-3."""
    m = hits["tri"] != phx.MISS
    pos = (rays["o"][m] + rays["d"][m] * hits["t"][m][:, None]).astype(np.float32)
    to_l = np.asarray(light, np.float32)[None, :] - pos
    dist = np.sqrt((to_l.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    out = np.zeros(len(pos), phx.RAY)
    out["o"] = pos
    out["d"] = (to_l / dist[:, None]).astype(np.float32)
    out["tmin"] = np.float32(1e-3)
    out["tmax"] = dist - np.float32(1e-3)
    return out


@dataclass
class RayPools:
    primary: np.ndarray
    primary_hits: np.ndarray
    bounce: np.ndarray   # valid bounce rays only (primary hits)
    shadow: np.ndarray


def ray_pools(ctx: phx.Context, scene: phx.Scene, width: int = 1920, height: int = 1080,
              frame: int = 0) -> RayPools:
    cam = camera(width, height)
    prim = ctx.gen_primary_rays(cam, width, height)
    hits = ctx.cast_closest(scene, prim)
    bounce = ctx.gen_bounce_rays(scene, prim, hits, BOUNCE_SEED, frame)
    valid = hits["tri"] != phx.MISS
    return RayPools(prim, hits, np.ascontiguousarray(bounce[valid]), shadow_rays(prim, hits))


@dataclass
class BatchSet:
    kinds: list        # "primary" | "diffuse" | "shadow"
    rays: list         # host arrays
    total: int = 0
    device: list = field(default_factory=list)  # (rays DeviceBuffer, out DeviceBuffer)


def batch_set(pools: RayPools, num_batches: int = 256, seed: int = BATCH_SEED,
              median: int = 32768, sigma: float = 1.5, lo: int = 1024, hi: int = 2097152,
              mix=(0.4, 0.4, 0.2)) -> BatchSet:
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(math.log(median) + sigma * rng.standard_normal(num_batches)), lo, hi)
    sizes = sizes.astype(np.int64)
    kinds = rng.choice(["primary", "diffuse", "shadow"], size=num_batches, p=list(mix))
    pool = {"primary": pools.primary, "diffuse": pools.bounce, "shadow": pools.shadow}
    rays = []
    for n, k in zip(sizes, kinds):
        p = pool[k]
        start = int(rng.integers(0, len(p)))
        idx = (start + np.arange(n)) % len(p)
        rays.append(np.ascontiguousarray(p[idx]))
    return BatchSet(list(kinds), rays, int(sizes.sum()))


def upload_batches(ctx: phx.Context, bs: BatchSet) -> None:
    for k, r in zip(bs.kinds, bs.rays):
        rb = ctx.alloc(len(r) * 32).upload(r)
        ob = ctx.alloc(len(r) * (1 if k == "shadow" else 16))
        bs.device.append((rb, ob))


def free_batches(bs: BatchSet) -> None:
    for rb, ob in bs.device:
        rb.free()
        ob.free()
    bs.device.clear()


def plan_for(ctx: phx.Context, scene: phx.Scene, bs: BatchSet, track: bool = True) -> phx.Plan:
    spec = []
    for k, r, (rb, ob) in zip(bs.kinds, bs.rays, bs.device):
        kind = phx.BATCH_ANYHIT if k == "shadow" else phx.BATCH_CLOSEST
        spec.append((scene, rb.ptr, ob.ptr, len(r), kind))
    return phx.Plan(ctx, spec, track=track)


def launch_plain(ctx: phx.Context, scene: phx.Scene, bs: BatchSet) -> None:
    """The per-batch form: one launch per batch on the context stream (asynchronous)."""
    for k, r, (rb, ob) in zip(bs.kinds, bs.rays, bs.device):
        if k == "shadow":
            ctx.cast_anyhit_device(scene, rb.ptr, len(r), ob.ptr, asynchronous=True)
        else:
            ctx.cast_closest_device(scene, rb.ptr, len(r), ob.ptr, asynchronous=True)


def download_outputs(bs: BatchSet) -> list:
    out = []
    for k, r, (rb, ob) in zip(bs.kinds, bs.rays, bs.device):
        out.append(ob.download(np.uint8 if k == "shadow" else phx.HIT, len(r)))
    return out
