"""The benchmark configurations at their stated sizes on the GPU, checked against the oracle.

  548,288-triangle scene   the 512x512 heightfield + 2,000 boxes as stated, pinned to the digest
                           of the reference's own castRay on it, plus its 1920x1080 primary,
                           diffuse-bounce and shadow rays against the oracle
  batched form             256 lognormal batches (1k..2M rays, 29.6 M rays, 40/40/20 mix) on the
                           1,004,000-triangle scene in one autotuned fused launch and its graph
                           replays: every batch equals its plain per-batch cast, and an evenly
                           strided 1/64 of every batch equals the oracle
  served form              4 resident scenes (100k, 1M, 4M, 8M triangles) under a 64 GB budget,
                           300 requests of 4k..4M rays: every request equals its plain cast alone,
                           and a strided 1/64 of every request equals the oracle
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from phantasyenginetestbed_amd import phx
from phantasyenginetestbed_amd import workloads as W

pytestmark = pytest.mark.gpu

STRIDE = 64


def say(msg):
    print(msg, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx():
    c = phx.Context(0)
    yield c
    c.destroy()


def oracle_out(bvh, kind, rays):
    if kind == "shadow":
        return O.cast_anyhit(bvh.nodes, bvh.tris, rays)
    return O.cast_closest(bvh.nodes, bvh.tris, rays)


@pytest.mark.timeout(900)
def test_heightfield512_as_stated(ctx, golden):
    g = golden["large"]["heightfield512"]
    say("[gpu-cfg] 512x512 heightfield + 2000 boxes")
    tris, datas, bvh = W.heightfield_scene(512, 2000)
    assert len(tris) == g["num_input"] == 548288
    assert (len(bvh.nodes), len(bvh.tris)) == (g["num_nodes"], g["num_refs"])
    assert sha(bvh.nodes) == g["nodes_sha256"] and sha(bvh.tris) == g["tris_sha256"]
    s = ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas)
    try:
        cam = W.camera(1920, 1080)
        small = ctx.gen_primary_rays(W.camera(480, 270), 480, 270)
        assert sha(small) == g["rays_480x270_sha256"]
        assert sha(ctx.cast_closest(s, small)) == g["hits_480x270_sha256"]
        say("[gpu-cfg] 548k scene: 1920x1080 primary, bounce and shadow rays")
        pools = W.ray_pools(ctx, s)
        assert pools.primary.tobytes() == O.primary_rays(cam.as_array(), 1920, 1080).tobytes()
        assert pools.primary_hits.tobytes() == O.cast_closest(bvh.nodes, bvh.tris,
                                                              pools.primary).tobytes()
        for kernel in (0, phx.KERNEL_ONE_RAY_PER_LANE, phx.KERNEL_PACKET):
            got = ctx.cast_closest(s, pools.bounce, kernel)
            if kernel == 0:
                want = O.cast_closest(bvh.nodes, bvh.tris, pools.bounce)
            assert got.tobytes() == want.tobytes(), f"bounce, kernel {kernel}"
            lit = ctx.cast_anyhit(s, pools.shadow, kernel)
            if kernel == 0:
                want_lit = O.cast_anyhit(bvh.nodes, bvh.tris, pools.shadow)
            assert (lit == want_lit).all(), f"shadow, kernel {kernel}"
    finally:
        s.release()


@pytest.mark.timeout(900)
def test_batched_form_full_size(ctx):
    say("[gpu-cfg] batched form: 1,004,000-triangle scene")
    _, _, bvh = W.heightfield_scene(700, 2000)
    s = ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas)
    bs = W.batch_set(W.ray_pools(ctx, s))
    assert len(bs.rays) == 256 and 25_000_000 < bs.total < 35_000_000
    sizes = np.array([len(r) for r in bs.rays])
    assert sizes.min() >= 1024 and sizes.max() <= 2097152
    say(f"[gpu-cfg] {bs.total} rays in 256 batches; uploading")
    W.upload_batches(ctx, bs)
    plan = None
    try:
        W.launch_plain(ctx, s, bs)
        ctx.synchronize()
        plain = W.download_outputs(bs)
        plan = W.plan_for(ctx, s, bs, track=True)
        say(f"[gpu-cfg] autotuned plan: {plan.autotune()}")
        for label, graph in (("fused", False), ("graph", True), ("graph replay", True)):
            for kind, rays, (_, ob) in zip(bs.kinds, bs.rays, bs.device):  # poison the outputs
                ob.upload(np.full(len(rays) * (1 if kind == "shadow" else 16), 0xAB, np.uint8))
            plan.launch(graph=graph)
            ctx.cast_closest(s, bs.rays[0][:4096])  # eager work between replays
            ctx.synchronize()
            got = W.download_outputs(bs)
            for i in range(256):
                assert got[i].tobytes() == plain[i].tobytes(), f"{label}: batch {i}"
            say(f"[gpu-cfg] {label}: 256 batches equal their plain casts")
        say("[gpu-cfg] strided 1/64 of every batch against the oracle")
        for i, (kind, rays) in enumerate(zip(bs.kinds, bs.rays)):
            sample = np.ascontiguousarray(rays[::STRIDE])
            want = oracle_out(bvh, kind, sample)
            assert plain[i][::STRIDE].tobytes() == want.tobytes(), f"batch {i} ({kind})"
    finally:
        if plan:
            plan.release()
        W.free_batches(bs)
        s.release()


@pytest.mark.timeout(900)
def test_served_form_full_size(ctx):
    scenes = []
    rs = None
    srv = None
    try:
        for name in ("100k", "1M", "4M", "8M"):
            say(f"[gpu-cfg] served scene {name}")
            scenes.append(W.served_scene(ctx, name, keep_bvh=True))
        assert [sc.triangles for sc in scenes] == [100000, 1004000, 4022792, 8024000]
        rs = W.request_set(len(scenes), 300)
        assert rs.sizes.min() >= 4096 and rs.sizes.max() <= 4194304
        W.allocate_outputs(rs)
        reqs = W.server_requests(scenes, rs)
        srv = phx.Server(ctx, 64 << 30, slots=3)
        for sc in scenes:
            srv.add_scene(sc.scene)
        for label, arrivals in (("burst", np.zeros(len(reqs))),
                                ("poisson", W.poisson_arrivals(len(reqs), 4.0, 3))):
            say(f"[gpu-cfg] served {rs.total} rays in {len(reqs)} requests ({label})")
            rs.out.array[:] = 0xAB
            lat, mk = srv.replay(reqs, arrivals)
            assert (lat > 0).all() and mk > 0
            mix = [W.request_output(rs, i) for i in range(len(reqs))]
            if label == "burst":
                first = mix
            else:
                for i in range(len(reqs)):
                    assert mix[i].tobytes() == first[i].tobytes(), f"request {i}: burst vs poisson"
        st = srv.stats()
        assert st["requestsDone"] == 2 * len(reqs) and st["bytesInUse"] == 0
        say("[gpu-cfg] every request alone, and a strided 1/64 against the oracle")
        for i in range(len(reqs)):
            sc = scenes[int(rs.scene_index[i])]
            kind = rs.kinds[i]
            o, n = int(rs.offsets[i]), int(rs.sizes[i])
            rays = sc.pools[kind].array[o:o + n]
            alone = ctx.cast_anyhit(sc.scene, rays) if kind == "shadow" else \
                ctx.cast_closest(sc.scene, rays)
            assert first[i].tobytes() == alone.tobytes(), f"request {i} ({kind}, {n}) in the mix"
            want = oracle_out(sc.bvh, kind, np.ascontiguousarray(rays[::STRIDE]))
            assert alone[::STRIDE].tobytes() == want.tobytes(), f"request {i} vs oracle"
    finally:
        if srv:
            srv.destroy()
        if rs is not None and rs.out is not None:
            rs.out.free()
        for sc in scenes:
            W.free_served_scene(sc)
