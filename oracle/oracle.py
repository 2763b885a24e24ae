"""ctypes binding of the oracle (liboracle.so) and of the reference's own CPU code compiled from
its sources (_ref/libphxref.so). TEST INFRASTRUCTURE ONLY: imported by tests/, by
__graft_entry__.smoke() to check the HIP library's results, and by bench.py's cpu_baseline leg.

Arrays use the numpy dtypes of phantasyenginetestbed_amd.phx (same byte layouts).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libphxref.so"
REF_FASTMATH_LIB = HERE / "_ref" / "libphxref_fastmath.so"
REFERENCE_ROOT = Path("/root/reference")

_P = C.c_void_p
_U32 = C.c_uint32


class Counters(C.Structure):
    _fields_ = [("nodesVisited", C.c_uint64), ("trianglesIntersected", C.c_uint64),
                ("triangleIntersectionTests", C.c_uint64), ("aabbIntersectionTests", C.c_uint64),
                ("aabbIntersections", C.c_uint64)]


def build(reference: bool | None = None) -> None:
    """Compiles liboracle.so, and _ref/libphxref.so when the reference tree is present."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "liboracle.so"], check=True)
    if reference is None:
        reference = REFERENCE_ROOT.exists()
    if reference:
        subprocess.run(["make", "-s", "-B", "-C", str(HERE), "ref", "asship"], check=True)


_olib = None
_rlib = {}


def olib() -> C.CDLL:
    global _olib
    if _olib is None:
        if not ORACLE_LIB.exists():
            build(reference=False)
        _olib = _configure(C.CDLL(str(ORACLE_LIB)))
    return _olib


@contextlib.contextmanager
def use_library(path):
    """Routes every oracle call of this module through another build of oracle.c + frame.c (a
    copy with a deliberate mistake planted, in the tests that show a check catches it)."""
    global _olib
    saved = olib()
    _olib = _configure(C.CDLL(str(path)))
    try:
        yield _olib
    finally:
        _olib = saved


def _configure(L: C.CDLL) -> C.CDLL:
    L.oracle_cast_closest.argtypes = [_P, _P, _P, _U32, _P, _U32]
    L.oracle_cast_debug.argtypes = [_P, _P, _P, _U32, _P, C.POINTER(Counters), _P, _U32]
    L.oracle_cast_anyhit.argtypes = [_P, _P, _P, _U32, _P, _U32]
    L.oracle_anyhit_debug.argtypes = [_P, _P, _P, _U32, _P, C.POINTER(Counters), _U32]
    L.oracle_brute_closest.argtypes = [_P, _U32, _P, _U32, _P]
    L.oracle_primary_rays.argtypes = [_P, _U32, _U32, C.c_float, C.c_float, _U32, _P]
    L.oracle_triangle.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P]
    L.oracle_triangle.restype = C.c_int
    L.oracle_aabb.argtypes = [_P, _P, _P, _P, _P, _P]
    L.oracle_aabb.restype = C.c_int
    F = C.c_float
    L.oracle_turn_sincos.argtypes = [F, C.POINTER(F), C.POINTER(F)]
    L.oracle_ggx.argtypes = [F, F]
    L.oracle_ggx.restype = F
    L.oracle_falloff.argtypes = [F, F]
    L.oracle_falloff.restype = F
    L.oracle_shade.argtypes = [_P, _P, _P, F, F, _P, F, _P, F, _P]
    L.oracle_sample_lambert.argtypes = [F, F, _P, _P]
    L.oracle_primary_rays_jittered.argtypes = [_P, _U32, _U32, F, F, F, F, _P]
    L.oracle_bounce_rays.argtypes = [_P, _P, _P, _U32, C.c_uint64, _U32, _P]
    L.oracle_frame.argtypes = [_P, _P, _P, _P, _U32, _P, _U32, _P, _U32, _P, _U32, _U32,
                               F, F, C.c_uint64, _U32, _P, _P, _P, _U32, _P, _U32]
    L.oracle_interpret.argtypes = [_P, _P, _U32, _P, _P, _U32, _P, _U32, _P]
    L.oracle_texel_coord.argtypes = [F, _U32]
    L.oracle_texel_coord.restype = _U32
    L.oracle_linear_table.argtypes = [_P]
    L.oracle_taa.argtypes = [_P, _P, _P, _P, _P, _U32, _U32, C.c_int, C.c_int, _P, _P]
    L.oracle_velocity.argtypes = [_P, _P, _P, _P, _U32, _U32, _P]
    L.oracle_halton23.argtypes = [_U32, C.POINTER(F), C.POINTER(F)]
    L.oracle_sample_ggx.argtypes = [F, F, F, _P, _P, _P]
    L.oracle_instance_prepare.argtypes = [_P, _P, _P, _P]
    L.oracle_instance_prepare.restype = C.c_int
    L.oracle_cast_instances.argtypes = [_P, _P, _P, _U32, _P, _U32, _P, _P, _P]
    L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, _U32, _P]
    return L


def have_ref(fastmath: bool = False) -> bool:
    return (REF_FASTMATH_LIB if fastmath else REF_LIB).exists()


def rlib(fastmath: bool = False) -> C.CDLL:
    if fastmath not in _rlib:
        path = REF_FASTMATH_LIB if fastmath else REF_LIB
        if not path.exists():
            raise RuntimeError(f"{path} not built (needs the reference tree)")
        L = C.CDLL(str(path))
        L.phxref_build.argtypes = [_P, _P, _U32, C.POINTER(_U32), C.POINTER(_U32)]
        L.phxref_build.restype = _P
        L.phxref_get.argtypes = [_P, _P, _P, _P]
        L.phxref_free.argtypes = [_P]
        L.phxref_cast.argtypes = [_P, _P, _P, _U32, _P, _U32]
        L.phxref_cast_bands.argtypes = [_P, _P, _P, _U32, _P, _U32]
        L.phxref_cast_debug.argtypes = [_P, _P, _P, _U32, _P, _P]
        L.phxref_primary_rays.argtypes = [_P, _P, _P, C.c_float, _U32, _U32, C.c_float,
                                          C.c_float, _P, _P]
        _rlib[fastmath] = L
    return _rlib[fastmath]


def _p(a):
    """The data pointer of a C-contiguous array (the C functions read it as packed records); a
    strided view would hand them the wrong memory, so it is refused."""
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("oracle inputs must be C-contiguous arrays (use np.ascontiguousarray)")
    return a.ctypes.data


# ---- oracle (C restatement) -----------------------------------------------------------------

def cast_closest(nodes, tris, rays, threads: int = 0):
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    olib().oracle_cast_closest(_p(nodes), _p(tris), _p(rays), len(rays), _p(hits), threads)
    return hits


def cast_debug(nodes, tris, rays, threads: int = 0, per_ray: bool = False):
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    tot = Counters()
    pr = np.empty((len(rays), 5), np.uint32) if per_ray else None
    olib().oracle_cast_debug(_p(nodes), _p(tris), _p(rays), len(rays), _p(hits), C.byref(tot),
                             _p(pr), threads)
    totals = {f: getattr(tot, f) for f, _ in Counters._fields_}
    return hits, totals, pr


def cast_anyhit(nodes, tris, rays, threads: int = 0):
    out = np.empty(len(rays), np.uint8)
    olib().oracle_cast_anyhit(_p(nodes), _p(tris), _p(rays), len(rays), _p(out), threads)
    return out


def anyhit_debug(nodes, tris, rays, threads: int = 0):
    """(inLight, summed counters of the any-hit traversal)."""
    out = np.empty(len(rays), np.uint8)
    tot = Counters()
    olib().oracle_anyhit_debug(_p(nodes), _p(tris), _p(rays), len(rays), _p(out), C.byref(tot),
                               threads)
    return out, {f: getattr(tot, f) for f, _ in Counters._fields_}


def brute_closest(tris, rays):
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    olib().oracle_brute_closest(_p(tris), len(tris), _p(rays), len(rays), _p(hits))
    return hits


def primary_rays(cam12, w, h, t_min=1e-4, t_max=None, order=0):
    from phantasyenginetestbed_amd.phx import FLT_MAX, RAY
    cam = np.ascontiguousarray(cam12, np.float32)
    rays = np.empty(w * h, RAY)
    olib().oracle_primary_rays(_p(cam), w, h, t_min, FLT_MAX if t_max is None else t_max,
                               order, _p(rays))
    return rays


def triangle(p0, p1, p2, o, d):
    a = [np.ascontiguousarray(x, np.float32) for x in (p0, p1, p2, o, d)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = olib().oracle_triangle(*[_p(x) for x in a], C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def aabb(o, d, mn, mx):
    a = [np.ascontiguousarray(x, np.float32) for x in (o, d, mn, mx)]
    ti, to = C.c_float(), C.c_float()
    hit = olib().oracle_aabb(*[_p(x) for x in a], C.byref(ti), C.byref(to))
    return bool(hit), ti.value, to.value


# ---- instances ---------------------------------------------------------------------------------

class OrInstance(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tris", C.c_void_p), ("objectToWorld", C.c_float * 12),
                ("datas", C.c_void_p)]


def _instances(instances):
    """[(nodes, tris, objectToWorld[, datas])] as an OrInstance array."""
    arr = (OrInstance * max(1, len(instances)))()
    for i, spec in enumerate(instances):
        nodes, tris, m = spec[:3]
        datas = spec[3] if len(spec) > 3 else None
        arr[i] = OrInstance(nodes.ctypes.data, tris.ctypes.data,
                            (C.c_float * 12)(*np.asarray(m, np.float32).ravel()),
                            None if datas is None else datas.ctypes.data)
    return arr


def instance_prepare(nodes, tris, m):
    """(world -> object rows (3, 4), world box min, max), or None for a singular transform."""
    arr = _instances([(nodes, tris, m)])
    a = np.zeros(12, np.float32)
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    ok = olib().oracle_instance_prepare(C.cast(arr, _P), _p(a), _p(mn), _p(mx))
    return (a.reshape(3, 4), mn, mx) if ok else None


def cast_instances(instances, rays, scene=None, anyhit: bool = False):
    """instances: [(nodes, tris, objectToWorld 3x4)]; scene: (nodes, tris) or None. Returns
    (hits, instance ids) or, with anyhit, inLight."""
    from phantasyenginetestbed_amd.phx import HIT
    arr = _instances(instances)
    rays = np.ascontiguousarray(rays)
    sn, st = (None, None) if scene is None else scene
    if anyhit:
        out = np.empty(len(rays), np.uint8)
        olib().oracle_cast_instances(_p(sn), _p(st), C.cast(arr, _P), len(instances), _p(rays),
                                     len(rays), None, None, _p(out))
        return out
    hits = np.empty(len(rays), HIT)
    ids = np.empty(len(rays), np.uint32)
    olib().oracle_cast_instances(_p(sn), _p(st), C.cast(arr, _P), len(instances), _p(rays),
                                 len(rays), _p(hits), _p(ids), None)
    return hits, ids


# ---- wavefront frame (frame.c) ------------------------------------------------------------

def turn_sincos(r: float):
    s, c = C.c_float(), C.c_float()
    olib().oracle_turn_sincos(r, C.byref(s), C.byref(c))
    return s.value, c.value


def ggx(n_dot_h: float, a: float) -> float:
    return olib().oracle_ggx(n_dot_h, a)


def falloff(d: float, rng: float) -> float:
    return olib().oracle_falloff(d, rng)


def shade(n, v, albedo, rough, metal, l, dist, strength, rng):
    a = [np.ascontiguousarray(x, np.float32) for x in (n, v, albedo, l, strength)]
    out = np.zeros(3, np.float32)
    olib().oracle_shade(_p(a[0]), _p(a[1]), _p(a[2]), rough, metal, _p(a[3]), dist, _p(a[4]),
                        rng, _p(out))
    return out


def sample_lambert(r1: float, r2: float, n):
    nn = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    olib().oracle_sample_lambert(r1, r2, _p(nn), _p(out))
    return out


def sample_ggx(r1: float, r2: float, a: float, refl, n):
    rr = np.ascontiguousarray(refl, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    olib().oracle_sample_ggx(r1, r2, a, _p(rr), _p(nn), _p(out))
    return out


def pcg32(init: int, seq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    olib().oracle_pcg32(init, seq, n, _p(out))
    return out


def primary_rays_jittered(cam12, w, h, jx, jy, t_min=1e-4, t_max=None):
    """Record order: 8x8 tiles when w and h are multiples of 8, else row-major."""
    from phantasyenginetestbed_amd.phx import FLT_MAX, RAY
    cam = np.ascontiguousarray(cam12, np.float32)
    rays = np.empty(w * h, RAY)
    olib().oracle_primary_rays_jittered(_p(cam), w, h, jx, jy, t_min,
                                        FLT_MAX if t_max is None else t_max, _p(rays))
    return rays


def bounce_rays(datas, rays, hits, seed: int, frame: int):
    from phantasyenginetestbed_amd.phx import RAY
    out = np.empty(len(rays), RAY)
    olib().oracle_bounce_rays(_p(datas), _p(rays), _p(hits), len(rays), seed, frame, _p(out))
    return out


def frame(nodes, tris, datas, materials, lights, cam12, w, h, jx, jy, seed, frame_index,
          threads: int = 0, textures=(), instances=()):
    """One frame: returns (colour (h, w, 4) float32, G-buffer records, bounce rays).
    instances: [(nodes, tris, objectToWorld, datas)] drawn besides the scene."""
    from phantasyenginetestbed_amd.phx import HITINFO, RAY, texture_arrays
    cam = np.ascontiguousarray(cam12, np.float32)
    color = np.zeros((h, w, 4), np.float32)
    gbuf = np.zeros(w * h, HITINFO)
    bounce = np.zeros((w // 2) * (h // 2), RAY)
    tex, keep = texture_arrays(textures)
    olib().oracle_frame(_p(nodes), _p(tris), _p(datas), _p(materials), len(materials),
                        C.cast(tex, _P), len(textures), _p(lights), len(lights), _p(cam), w, h,
                        jx, jy, seed, frame_index, _p(color), _p(gbuf), _p(bounce), threads,
                        C.cast(_instances(instances), _P) if instances else None, len(instances))
    del keep
    return color, gbuf, bounce


def interpret(datas, materials, rays, hits, textures=()):
    """RayHitInfo records of closest hits (interpretRayHitKernel with material textures)."""
    from phantasyenginetestbed_amd.phx import HITINFO, texture_arrays
    out = np.zeros(len(rays), HITINFO)
    tex, keep = texture_arrays(textures)
    nm = 0 if materials is None else len(materials)
    olib().oracle_interpret(_p(rays), _p(hits), len(rays), _p(datas), _p(materials), nm,
                            C.cast(tex, _P), len(textures), _p(out))
    del keep
    return out


def texel_coord(c: float, dim: int) -> int:
    return olib().oracle_texel_coord(c, dim)


def linear_table() -> np.ndarray:
    out = np.zeros(256, np.float32)
    olib().oracle_linear_table(_p(out))
    return out


def taa(cur, hist_in, roughness, reset: bool, clamp: bool = True, vel_in=None, vel=None):
    """Returns (history out, output), both (h, w, 4) float32. vel / vel_in: this and the previous
    frame's velocity, (h, w, 2) pixels (None: zero)."""
    h, w = cur.shape[:2]
    cur = np.ascontiguousarray(cur, np.float32)
    hist_in = np.ascontiguousarray(hist_in if hist_in is not None else np.zeros_like(cur),
                                   np.float32)
    rough = None if roughness is None else np.ascontiguousarray(roughness, np.float32)
    vel_in = None if vel_in is None else np.ascontiguousarray(vel_in, np.float32)
    vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
    hist_out = np.empty_like(cur)
    out = np.empty_like(cur)
    olib().oracle_taa(_p(cur), _p(hist_in), _p(rough), _p(vel_in), _p(vel), w, h, int(reset),
                      int(clamp), _p(hist_out), _p(out))
    return hist_out, out


def velocity(cam12, prev_cam12, gbuf, prim, w: int, h: int) -> np.ndarray:
    """(h, w, 2) screen-space velocity in pixels (current minus previous camera's projection)."""
    cam = np.ascontiguousarray(cam12, np.float32)
    prev = np.ascontiguousarray(prev_cam12, np.float32)
    vel = np.zeros((h, w, 2), np.float32)
    olib().oracle_velocity(_p(cam), _p(prev), _p(gbuf), _p(prim), w, h, _p(vel))
    return vel


def halton23(f: int):
    jx, jy = C.c_float(), C.c_float()
    olib().oracle_halton23(f, C.byref(jx), C.byref(jy))
    return jx.value, jy.value


def pixel_ids(w: int, h: int) -> np.ndarray:
    """(h, w) array of record ids: 8x8 tiles when w, h are multiples of 8, else row-major."""
    y, x = np.mgrid[0:h, 0:w].astype(np.int64)
    if w % 8 == 0 and h % 8 == 0:
        return ((y // 8) * (w // 8) + x // 8) * 64 + (y % 8) * 8 + x % 8
    return y * w + x


# ---- the reference's own code ---------------------------------------------------------------

def ref_build(tris, datas=None, fastmath: bool = False):
    from phantasyenginetestbed_amd.phx import NODE, TRI, TRIDATA
    L = rlib(fastmath)
    nn, nt = C.c_uint32(), C.c_uint32()
    h = L.phxref_build(_p(tris), _p(datas), len(tris), C.byref(nn), C.byref(nt))
    nodes = np.empty(nn.value, NODE)
    out_tris = np.empty(nt.value, TRI)
    out_datas = np.empty(nt.value, TRIDATA)
    L.phxref_get(h, _p(nodes), _p(out_tris), _p(out_datas))
    L.phxref_free(h)
    return nodes, out_tris, out_datas


def ref_cast(nodes, tris, rays, threads: int = 0, fastmath: bool = False):
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    rlib(fastmath).phxref_cast(_p(nodes), _p(tris), _p(rays), len(rays), _p(hits), threads)
    return hits


def ref_cast_bands(nodes, tris, rays, threads: int = 10, fastmath: bool = False):
    """castRay over `threads` fixed contiguous bands, the reference CPU renderer's split
    (cpu_tracer_renderer/CPUTracerRenderer.cpp:77-87)."""
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    rlib(fastmath).phxref_cast_bands(_p(nodes), _p(tris), _p(rays), len(rays), _p(hits), threads)
    return hits


def ref_cast_debug(nodes, tris, rays):
    from phantasyenginetestbed_amd.phx import HIT
    hits = np.empty(len(rays), HIT)
    cnt = np.empty((len(rays), 5), np.uint32)
    rlib().phxref_cast_debug(_p(nodes), _p(tris), _p(rays), len(rays), _p(hits), _p(cnt))
    return hits, cnt


def ref_primary_rays(pos, dir, up, fov_rad, w, h, t_min, t_max):
    from phantasyenginetestbed_amd.phx import RAY
    rays = np.empty(w * h, RAY)
    cam = np.empty(12, np.float32)
    f3 = [np.ascontiguousarray(x, np.float32) for x in (pos, dir, up)]
    rlib().phxref_primary_rays(*[_p(x) for x in f3], fov_rad, w, h, t_min, t_max, _p(rays),
                               _p(cam))
    return rays, cam
