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


# ---- served form --------------------------------------------------------------------------------
#
#   4 resident scenes: 100k random triangles (seed 0xC0FFEE01), the 1,004,000-triangle heightfield,
#   a 1414x1414 heightfield + 2,000 boxes (4,022,792 triangles) and a 2000x2000 heightfield + 2,000
#   boxes (8,024,000 triangles); per scene ray pools of primary, diffuse-bounce and shadow rays, repeated to
#   4,194,304 rays and held in pinned host memory. Requests: sizes lognormal(ln 32768, 1.5) clipped
#   to [4096, 4194304], kinds 40% primary / 40% diffuse / 20% shadow, scene uniform (seed
#   0xC0FFEE05), each a contiguous slice of its pool.

SERVE_SEED = 0xC0FFEE05
RANDOM_CAM = ((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0))
RANDOM_LIGHT = (0.5, 2.0, 0.5)
POOL_RAYS = 4194304


@dataclass
class ServedScene:
    name: str
    triangles: int
    scene: object            # phx.Scene
    pools: dict              # kind -> phx.HostBuffer of RAY (POOL_RAYS records)
    bvh: object = None       # phx.BVH (host arrays), kept on request


def _scene_geometry(name: str):
    if name == "100k":
        t, d = phx.random_triangles(100000, 0xC0FFEE01)
        return t, d, RANDOM_CAM, RANDOM_LIGHT, (1024, 1024)
    if name == "1M":
        t, d = phx.heightfield_boxes(700, 2000, SCENE_SEED)
        return t, d, CAMERA, LIGHT, (1920, 1080)
    if name == "4M":
        t, d = phx.heightfield_boxes(1414, 2000, SCENE_SEED)
        return t, d, CAMERA, LIGHT, (1920, 1080)
    if name == "8M":
        t, d = phx.heightfield_boxes(2000, 2000, SCENE_SEED)
        return t, d, CAMERA, LIGHT, (1920, 1080)
    raise ValueError(name)


def _tile(rays: np.ndarray, n: int, dst: np.ndarray) -> None:
    reps = -(-n // len(rays))
    dst[:] = np.tile(rays, reps)[:n]


def served_scene(ctx: phx.Context, name: str, pool_rays: int = POOL_RAYS,
                 keep_bvh: bool = False) -> ServedScene:
    t, d, cam_spec, light, (w, h) = _scene_geometry(name)
    bvh = phx.BVH(t, d)
    scene = ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas)
    cam = phx.camera_look_at(*cam_spec, math.radians(FOV_DEG), w, h)
    prim = ctx.gen_primary_rays(cam, w, h)
    hits = ctx.cast_closest(scene, prim)
    valid = hits["tri"] != phx.MISS
    bounce = ctx.gen_bounce_rays(scene, prim, hits, BOUNCE_SEED, 0)[valid]
    shadow = shadow_rays(prim, hits, light)
    pools = {}
    for kind, rays in (("primary", prim), ("diffuse", bounce), ("shadow", shadow)):
        hb = phx.HostBuffer(phx.RAY, pool_rays)
        _tile(rays, pool_rays, hb.array)
        pools[kind] = hb
    return ServedScene(name, len(t), scene, pools, bvh if keep_bvh else None)


def free_served_scene(s: ServedScene) -> None:
    for hb in s.pools.values():
        hb.free()
    s.scene.release()


@dataclass
class RequestSet:
    scene_index: np.ndarray
    kinds: list
    offsets: np.ndarray
    sizes: np.ndarray
    out: object = None          # phx.HostBuffer (uint8) holding every request's output
    out_offsets: np.ndarray = None

    @property
    def total(self) -> int:
        return int(self.sizes.sum())


def request_set(num_scenes: int, n: int, seed: int = SERVE_SEED, median: int = 32768,
                sigma: float = 1.5, lo: int = 4096, hi: int = 4194304, mix=(0.4, 0.4, 0.2),
                pool_rays: int = POOL_RAYS) -> RequestSet:
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(math.log(median) + sigma * rng.standard_normal(n)), lo, hi)
    sizes = np.minimum(sizes.astype(np.int64), pool_rays)
    kinds = list(rng.choice(["primary", "diffuse", "shadow"], size=n, p=list(mix)))
    scene_index = rng.integers(0, num_scenes, n)
    offsets = np.array([int(rng.integers(0, pool_rays - s + 1)) for s in sizes], np.int64)
    return RequestSet(scene_index, kinds, offsets, sizes)


def allocate_outputs(rs: RequestSet) -> None:
    per = np.array([1 if k == "shadow" else 16 for k in rs.kinds], np.int64)
    nbytes = (rs.sizes * per + 255) // 256 * 256
    rs.out_offsets = np.concatenate([[0], np.cumsum(nbytes)[:-1]])
    rs.out = phx.HostBuffer(np.uint8, int(nbytes.sum()))


def server_requests(scenes: list, rs: RequestSet) -> list:
    reqs = []
    for i in range(len(rs.sizes)):
        s = scenes[int(rs.scene_index[i])]
        pool = s.pools[rs.kinds[i]]
        kind = phx.BATCH_ANYHIT if rs.kinds[i] == "shadow" else phx.BATCH_CLOSEST
        reqs.append(phx.Server.request(s.scene, pool.ptr + int(rs.offsets[i]) * 32,
                                       rs.out.ptr + int(rs.out_offsets[i]), int(rs.sizes[i]),
                                       kind))
    return reqs


def server_requests_device(ctx: phx.Context, scenes: list, rs: RequestSet):
    """The same requests with device-resident rays and outputs (PHX_MEM_DEVICE): every scene's
    ray pools copied to the device once, one device buffer for all outputs. Returns (requests,
    buffers to free, output buffer)."""
    pools = {}
    keep = []
    for si, s in enumerate(scenes):
        for kind, hb in s.pools.items():
            db = ctx.alloc(hb.array.nbytes)
            ctx.copy_to_device(db.ptr, hb.ptr, hb.array.nbytes)
            pools[(si, kind)] = db
            keep.append(db)
    per = np.array([1 if k == "shadow" else 16 for k in rs.kinds], np.int64)
    nbytes = (rs.sizes * per + 255) // 256 * 256
    offs = np.concatenate([[0], np.cumsum(nbytes)[:-1]])
    out = ctx.alloc(int(nbytes.sum()))
    keep.append(out)
    reqs = []
    for i in range(len(rs.sizes)):
        si = int(rs.scene_index[i])
        kind = phx.BATCH_ANYHIT if rs.kinds[i] == "shadow" else phx.BATCH_CLOSEST
        reqs.append(phx.Server.request(scenes[si].scene,
                                       pools[(si, rs.kinds[i])].ptr + int(rs.offsets[i]) * 32,
                                       out.ptr + int(offs[i]), int(rs.sizes[i]), kind,
                                       device=True))
    return reqs, keep, out


def request_output(rs: RequestSet, i: int) -> np.ndarray:
    n = int(rs.sizes[i])
    o = int(rs.out_offsets[i])
    if rs.kinds[i] == "shadow":
        return rs.out.array[o:o + n].copy()
    return rs.out.array[o:o + 16 * n].view(phx.HIT).copy()


def poisson_arrivals(n: int, rate_per_ms: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_ms, n))


INSTANCE_SEED = 0xC0FFEE05


@dataclass
class InstanceWorkload:
    meshes: list       # phx.Scene per mesh
    instances: list    # (phx.Scene, 3x4 object -> world matrix)
    rays: np.ndarray


def instance_workload(ctx: phx.Context, count: int = 300, num_rays: int = 1 << 20,
                      seed: int = INSTANCE_SEED) -> InstanceWorkload:
    """Dynamic-mesh workload: three random-triangle meshes (300 / 4,000 / 20,000 triangles, edges
    0.01-0.1) placed as `count` instances (a random rotation from a unit quaternion, scale 0.5-2
    per axis, every 7th instance mirrored, positions uniform in [-6, 6]^3), and `num_rays` rays
    from origins uniform in [-9, 9]^3 towards points uniform in [-6, 6]^3 (tMin 1e-4)."""
    rng = np.random.default_rng(seed)
    meshes = []
    for k, n in enumerate((300, 4000, 20000)):
        t, d = phx.random_triangles(n, seed + k, 0.01, 0.1)
        b = phx.BVH(t, d)
        meshes.append(ctx.upload_scene(b.nodes, b.tris, b.datas))
    instances = []
    for k in range(count):
        q = rng.normal(size=4)
        w, x, y, z = q / np.linalg.norm(q)
        rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        scale = rng.uniform(0.5, 2.0, 3)
        if k % 7 == 3:
            scale[0] = -scale[0]
        m = np.zeros((3, 4))
        m[:, :3] = rot @ np.diag(scale)
        m[:, 3] = rng.uniform(-6, 6, 3)
        instances.append((meshes[k % 3], m.astype(np.float32)))
    rays = np.zeros(num_rays, phx.RAY)
    rays["o"] = rng.uniform(-9, 9, (num_rays, 3))
    d = rng.uniform(-6, 6, (num_rays, 3)) - rays["o"]
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"] = np.float32(1e-4)
    rays["tmax"] = np.float32(phx.FLT_MAX)
    return InstanceWorkload(meshes, instances, rays)
