"""Handle lifetimes and request control through the C ABI: cancellation of served requests, scenes
released while a plan or server still uses them, a frame graph after a scene is released and
another uploaded, out-of-range triangle indices in caller-supplied hits, and oversized requests.
Every case is forced from the host; no kernel is made to fault."""
import time

import numpy as np
import pytest

from oracle import oracle as O
from phantasyenginetestbed_amd import phx
from phantasyenginetestbed_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = phx.Context(0)
    yield c
    c.destroy()


def heightfield(ctx, grid, boxes, materials=0, lights=0):
    tris, datas = phx.heightfield_boxes(grid, boxes, 0xC0FFEE02, num_materials=materials)
    bvh = phx.BVH(tris, datas)
    mats = phx.materials_random(materials, 5) if materials else None
    lts = phx.lights_random(lights, 5) if lights else None
    return bvh, mats, lts, ctx.upload_scene(bvh.nodes, bvh.tris, bvh.datas, mats, lts)


@pytest.fixture(scope="module")
def scene(ctx):
    bvh, _, _, s = heightfield(ctx, 48, 30)
    cam = W.camera(160, 100)
    prim = ctx.gen_primary_rays(cam, 160, 100)
    yield bvh, s, prim
    s.release()


def test_cancel_queued_and_admitted_requests(ctx, scene):
    bvh, s, prim = scene
    n = 8000
    rays = np.ascontiguousarray(prim[len(prim) - n:])
    want = O.cast_closest(bvh.nodes, bvh.tris, rays)
    per_req = n * 32 + n * 16  # rays + hits (both multiples of the arena's 256-B ranges)
    srv = phx.Server(ctx, s.device_bytes() + 2 * per_req + 1024, slots=2)
    bufs = []
    try:
        srv.add_scene(s)
        srv.pause(True)
        tickets = []
        for _ in range(5):
            rb, ob = phx.HostBuffer(phx.RAY, n), phx.HostBuffer(phx.HIT, n)
            rb.array[:] = rays
            ob.array.view(np.uint8)[:] = 0xAB
            bufs.append((rb, ob))
            tickets.append(srv.submit(phx.Server.request(s, rb.ptr, ob.ptr, n,
                                                         phx.BATCH_CLOSEST)))
        say("[gpu-life] paused server: 2 requests admitted, 3 queued")
        t0 = time.monotonic()
        while srv.stats()["bytesInUse"] < 2 * per_req and time.monotonic() - t0 < 5:
            time.sleep(0.01)
        st = srv.stats()
        assert st["queued"] == 5 and st["launches"] == 0 and st["bytesInUse"] == 2 * per_req
        assert srv.cancel(tickets[0])   # admitted: its arena space returns, request 2 is admitted
        assert srv.cancel(tickets[3])   # queued
        assert srv.cancel(tickets[3])   # cancelling twice reports it cancelled again
        with pytest.raises(phx.PhxError) as ei:
            srv.poll(tickets[0])
        assert ei.value.status == phx.ERR_CANCELLED
        t0 = time.monotonic()  # the freed space admits the third request
        while srv.stats()["bytesInUse"] < 2 * per_req and time.monotonic() - t0 < 5:
            time.sleep(0.01)
        st = srv.stats()
        assert st["cancelled"] == 2 and st["queued"] == 3 and st["bytesInUse"] == 2 * per_req
        srv.pause(False)
        for i in (1, 2, 4):
            assert srv.wait(tickets[i]) > 0
            assert bufs[i][1].array.tobytes() == want.tobytes(), f"request {i}"
        for i in (0, 3):
            with pytest.raises(phx.PhxError) as ei:
                srv.wait(tickets[i])
            assert ei.value.status == phx.ERR_CANCELLED
            assert (bufs[i][1].array.view(np.uint8) == 0xAB).all(), "cancelled output written"
        st = srv.stats()
        assert st["requestsDone"] == 3 and st["cancelled"] == 2 and st["bytesInUse"] == 0
        # a completed request can no longer be cancelled
        t = srv.submit(phx.Server.request(s, bufs[1][0].ptr, bufs[1][1].ptr, n,
                                          phx.BATCH_CLOSEST))
        t0 = time.monotonic()
        while not srv.poll(t)[0] and time.monotonic() - t0 < 30:
            time.sleep(0.001)
        assert not srv.cancel(t)
        srv.wait(t)
        with pytest.raises(phx.PhxError):
            srv.cancel(t)  # forgotten ticket
    finally:
        srv.destroy()
        for rb, ob in bufs:
            rb.free()
            ob.free()


def test_oversized_request_is_refused(ctx, scene):
    _, s, _ = scene
    srv = phx.Server(ctx, 1 << 30, slots=1)
    try:
        srv.add_scene(s)
        dev = ctx.alloc(256)
        try:
            with pytest.raises(phx.PhxError) as ei:  # refused before any buffer is touched
                srv.submit(phx.Server.request(s, dev.ptr, dev.ptr, 0x80000000, phx.BATCH_CLOSEST,
                                              device=True))
            assert ei.value.status == phx.ERR_INVALID_ARGUMENT
        finally:
            dev.free()
    finally:
        srv.destroy()


def test_released_scene_stays_valid_for_its_plan_and_server(ctx):
    say("[gpu-life] release a scene held by a plan and a server")
    bvh, _, _, s = heightfield(ctx, 40, 20)
    rays = ctx.gen_primary_rays(W.camera(96, 64), 96, 64)
    want = O.cast_closest(bvh.nodes, bvh.tris, rays)
    rbuf = ctx.alloc(len(rays) * 32).upload(rays)
    hbuf = ctx.alloc(len(rays) * 16)
    plan = phx.Plan(ctx, [(s, rbuf.ptr, hbuf.ptr, len(rays), phx.BATCH_CLOSEST)])
    srv = phx.Server(ctx, 1 << 28, slots=1)
    srv.add_scene(s)
    handle = s.handle
    s.release()
    try:
        plan.launch()
        assert hbuf.download(phx.HIT, len(rays)).tobytes() == want.tobytes()
        with pytest.raises(phx.PhxError) as ei:  # the released handle is refused
            srv.submit(phx.Request(handle, rbuf.ptr, hbuf.ptr, len(rays), phx.BATCH_CLOSEST,
                                   phx.MEM_DEVICE, 0))
        assert ei.value.status == phx.ERR_INVALID_HANDLE
        srv.destroy()
        plan.launch(graph=True)  # still held by the plan
        assert hbuf.download(phx.HIT, len(rays)).tobytes() == want.tobytes()
    finally:
        srv.destroy()
        plan.release()  # the last hold: the scene's arrays are freed here
        rbuf.free()
        hbuf.free()



def test_context_destroyed_before_its_objects(ctx):
    """phx_destroy while a scene, plan, renderer, instance set and server made from the context
    still exist: the handle is refused at once (so is a plan launch, which takes it), the server
    still serves with the right hits, and releasing the objects afterwards, scene first, frees
    the rest."""
    say("[gpu-life] destroy a context before its objects")
    c = phx.Context(0)
    bvh, _, _, s = heightfield(c, 40, 20)
    rays = c.gen_primary_rays(W.camera(96, 64), 96, 64)
    want = O.cast_closest(bvh.nodes, bvh.tris, rays)
    # buffers of the long-lived context, readable after `c` is gone
    rbuf = ctx.alloc(len(rays) * 32).upload(rays)
    hbuf = ctx.alloc(len(rays) * 16)
    sbuf = ctx.alloc(len(rays) * 16)
    plan = phx.Plan(c, [(s, rbuf.ptr, hbuf.ptr, len(rays), phx.BATCH_CLOSEST)])
    plan.launch()
    assert hbuf.download(phx.HIT, len(rays)).tobytes() == want.tobytes()
    r = c.renderer(64, 48)
    iset = phx.InstanceSet(c, [(s, np.eye(3, 4, dtype=np.float32))])
    srv = phx.Server(c, 1 << 28, slots=2)
    srv.add_scene(s)
    handle = c.handle
    assert phx.lib().phx_context_num_cus(handle) > 0
    c.destroy()
    try:
        assert phx.lib().phx_context_num_cus(handle) == 0  # the handle is gone at once
        with pytest.raises(phx.PhxError) as ei:
            plan.launch(graph=True)
        assert ei.value.status == phx.ERR_INVALID_HANDLE
        t = srv.submit(phx.Request(s.handle, rbuf.ptr, sbuf.ptr, len(rays), phx.BATCH_CLOSEST,
                                   phx.MEM_DEVICE, 0))
        srv.wait(t)
        assert sbuf.download(phx.HIT, len(rays)).tobytes() == want.tobytes()
    finally:
        # any order: the scene first (still held by the plan, the set and the server)
        s.release()
        srv.destroy()
        r.release()
        plan.release()
        iset.release()  # the last holder: the context's stream and buffers go here
        for b in (rbuf, hbuf, sbuf):
            b.free()

def test_frame_graph_recaptures_for_a_new_scene(ctx):
    """Release a scene, upload another (often at the same host address): a renderer replaying
    captured frames must capture again and render the new scene exactly as a direct frame."""
    w, h = 48, 32
    cam = W.camera(w, h)
    graph, direct = ctx.renderer(w, h), ctx.renderer(w, h)
    try:
        bvh, mats, lts, a = heightfield(ctx, 40, 20, materials=8, lights=3)
        for f in range(3):
            graph.render(a, cam, f, 7, phx.FRAME_GRAPH)
        a.release()
        for grid, boxes in ((56, 40), (40, 20)):
            say(f"[gpu-life] frame graph after re-upload ({grid}x{grid} + {boxes} boxes)")
            bvh, mats, lts, b = heightfield(ctx, grid, boxes, materials=8, lights=2)
            try:
                for f in range(3, 6):
                    got = graph.render(b, cam, f, 7, phx.FRAME_GRAPH)
                    want = direct.render(b, cam, f, 7, 0)
                    assert got.tobytes() == want.tobytes(), f"frame {f}"
                ref, _, _ = O.frame(bvh.nodes, bvh.tris, bvh.datas, mats, lts, cam.as_array(), w,
                                    h, 0.0, 0.0, 7, 5)
                assert got.tobytes() == ref.tobytes()
            finally:
                b.release()
    finally:
        graph.release()
        direct.release()


def test_hits_of_another_scene_read_as_misses(ctx, scene):
    bvh, s, prim = scene
    rays = np.ascontiguousarray(prim[len(prim) // 2:len(prim) // 2 + 64])
    hits = O.cast_closest(bvh.nodes, bvh.tris, rays)
    assert (hits["tri"] != phx.MISS).sum() > 10
    bad = hits.copy()
    bad["tri"][::3] = len(bvh.tris)        # one past the last triangle
    bad["tri"][1::3] = 0x7FFFFFFF          # far outside
    out = ctx.gen_bounce_rays(s, rays, bad, 3, 0)
    ref = O.bounce_rays(bvh.datas, rays, hits, 3, 0)
    keep = np.arange(64) % 3 == 2
    assert out[keep].tobytes() == ref[keep].tobytes()
    dummy = out[~keep]
    assert (dummy["o"] == 1e6).all() and (dummy["tmin"] == np.float32(0.001)).all()
    assert (dummy["tmax"] == np.float32(0.001)).all()
    info = ctx.interpret_hits(s, rays, bad)
    assert (info["alpha"][~keep] == -1).all()
    assert (info["pos"][~keep] == 0).all()
    assert info[keep].tobytes() == O.interpret(bvh.datas, None, rays[keep],
                                               hits[keep]).tobytes()
