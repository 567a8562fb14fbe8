"""GPU: the serving form (rt_server_*): device camera rays, continuous batching, and every
result identical alone and in a mix of concurrent requests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(rt, scenes, ctx):
    return [rt.Scene(ctx, scenes.torus_scene(), desc="rayhierarchy"),
            rt.Scene(ctx, scenes.random_triangles(1024, seed=1), desc="rayhierarchy"),
            rt.Scene(ctx, scenes.cluster_scene(20000, seed=11), desc="rayhierarchy")]


def test_request_matches_plain_trace(rt, scenes, gpu_ctx):
    """spp = 1: the device camera reproduces the numpy pinhole camera bit for bit, so a
    served request equals the plain per-scene traces of the same rays."""
    scs = _scenes(rt, scenes, gpu_ctx)
    srv = rt.Server(gpu_ctx, scs, max_active=8)
    try:
        t = srv.submit(0, 128, 96, 1, origin=(0.0, 0.0, -3.0), fov_deg=60.0,
                       light=(2.0, 2.0, -2.0), outputs=True)
        res = srv.wait(t, timeout_s=60.0)
    finally:
        srv.close()
    rays = scenes.pinhole_rays(128, 96)
    d = rt.DeviceRays.from_numpy(rays)
    h = scs[0].trace_closest(d, rt.RT_CLOSEST_PRIMARY)
    ids, ts = h.id.cpu().numpy(), h.t.cpu().numpy()
    print(f"[serve] hits {int((ids >= 0).sum())} / {ids.size}", flush=True)
    assert np.array_equal(res["id"], ids)
    assert np.array_equal(res["t"].view(np.uint32), ts.view(np.uint32))
    assert res["shadow_rays"] == int((ids >= 0).sum())
    sh = rt.make_shadow_rays(d, h, (2.0, 2.0, -2.0))
    occ = scs[0].trace_occluded(sh).cpu().numpy()
    assert np.array_equal(res["occ"], occ)
    for s in scs:
        s.close()


def test_alone_equals_mix(rt, scenes, gpu_ctx):
    scs = _scenes(rt, scenes, gpu_ctx)
    g = np.random.default_rng(5)
    reqs = []
    for k in range(16):
        reqs.append(dict(scene=int(g.integers(0, 3)), width=int(g.choice([32, 64, 128])),
                         height=int(g.choice([32, 64, 96])), spp=int(g.integers(1, 4)),
                         origin=(0.0, 0.0, -12.0 if k % 2 else -3.0),
                         fov_deg=50.0 if k % 2 else 60.0, light=(2.0, 2.0, -2.0), seed=k))
    srv = rt.Server(gpu_ctx, scs, max_active=64)
    try:
        alone = []
        for q in reqs:  # one at a time
            alone.append(srv.wait(srv.submit(**q, outputs=True), timeout_s=60.0))
        t0, m0 = srv.counters()
        tickets = [srv.submit(**q, outputs=True) for q in reqs]  # all at once
        mixed = [srv.wait(t, timeout_s=120.0) for t in tickets]
        t1, m1 = srv.counters()
    finally:
        srv.close()
    print(f"[serve] alone ticks {t0}, mix ticks {t1 - t0}, largest tick {m1} jobs", flush=True)
    assert m1 > 2, "the mix was not batched"
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a["id"], b["id"]), f"request {k}: ids differ alone vs mix"
        assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32)), f"request {k} t"
        assert np.array_equal(a["occ"], b["occ"]), f"request {k}: occlusion differs"
    for s in scs:
        s.close()


def test_budget_admission(rt, scenes, gpu_ctx):
    """A budget smaller than two requests admits them one by one; all complete."""
    scs = _scenes(rt, scenes, gpu_ctx)
    srv = rt.Server(gpu_ctx, scs, mem_budget=1 << 20, max_active=64)
    try:
        tickets = [srv.submit(1, 64, 64, 1, outputs=False) for _ in range(6)]
        stats = [srv.wait(t, timeout_s=60.0) for t in tickets]
        ticks, most = srv.counters()
    finally:
        srv.close()
    assert all(s["done_s"] >= s["admit_s"] >= s["submit_s"] for s in stats)
    assert most <= 2, f"{most} jobs in one tick despite the budget"
    print(f"[serve] budget run: {ticks} ticks", flush=True)
    for s in scs:
        s.close()


def test_bad_request_fails_alone(rt, scenes, gpu_ctx):
    """A request whose buffers exceed the server's pool is refused at submit (RT_ERR_OOM) and
    touches no other request; a naive scene is refused at server creation.  A scene of
    16x16-ray leaves (4,096-ray blocks, per-ray path) is served like any other."""
    good = rt.Scene(gpu_ctx, scenes.random_triangles(1024, seed=1), desc="rayhierarchy")
    big = rt.Scene(gpu_ctx, scenes.random_triangles(1024, seed=1),
                   desc=rt.default_desc("rayhierarchy", chunk_x=16, chunk_y=16))
    naive = rt.Scene(gpu_ctx, scenes.random_triangles(64, seed=2), desc="naive")
    with pytest.raises(rt.RtError):
        rt.Server(gpu_ctx, [good, naive])
    srv = rt.Server(gpu_ctx, [good, big], mem_budget=32 << 20, max_active=8)
    try:
        t = srv.submit(0, 64, 64, 1, outputs=True)
        with pytest.raises(rt.RtError) as e:
            srv.submit(1, 2048, 2048, 4)
        assert e.value.status == 3  # RT_ERR_OOM
        t_big = srv.submit(1, 64, 64, 1, outputs=True)
        res = srv.wait(t, timeout_s=60.0)
        res_big = srv.wait(t_big, timeout_s=60.0)
    finally:
        srv.close()
    ids_big = big.trace_closest(rt.DeviceRays.from_numpy(scenes.pinhole_rays(64, 64)))
    assert np.array_equal(res_big["id"], ids_big.id.cpu().numpy())
    rays = scenes.pinhole_rays(64, 64)
    ids = good.trace_closest(rt.DeviceRays.from_numpy(rays)).id.cpu().numpy()
    assert np.array_equal(res["id"], ids)
    for s in (good, big, naive):
        s.close()


def test_raybvh_request_matches_plain_trace(rt, scenes, gpu_ctx):
    """ray-bvh scenes are served like rayhierarchy ones: a request equals the plain traces."""
    v9 = scenes.room_scene(nclusters=40)
    sc = rt.Scene(gpu_ctx, v9, desc="ray-bvh")
    srv = rt.Server(gpu_ctx, [sc], max_active=8)
    try:
        t = srv.submit(0, 96, 64, 1, origin=(0.0, 0.0, -9.0), fov_deg=50.0,
                       light=(1.0, 2.5, -8.0), outputs=True)
        res = srv.wait(t, timeout_s=60.0)
    finally:
        srv.close()
    rays = scenes.pinhole_rays(96, 64, origin=(0.0, 0.0, -9.0), fov_deg=50.0)
    d = rt.DeviceRays.from_numpy(rays)
    h = sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY)
    assert np.array_equal(res["id"], h.id.cpu().numpy())
    sh = rt.make_shadow_rays(d, h, (1.0, 2.5, -8.0))
    assert np.array_equal(res["occ"], sc.trace_occluded(sh).cpu().numpy())
    sc.close()


def test_pipelined_ticks_many_requests(rt, scenes, gpu_ctx):
    """Requests arriving over many ticks (two ticks in flight at once, blocks reused from the
    pool) each equal the same request traced alone."""
    scs = _scenes(rt, scenes, gpu_ctx)
    g = np.random.default_rng(21)
    reqs = [dict(scene=int(g.integers(0, 3)), width=int(g.choice([16, 48, 80])),
                 height=int(g.choice([16, 32, 64])), spp=int(g.integers(1, 3)),
                 origin=(0.0, 0.0, -12.0), fov_deg=50.0, light=(2.0, 2.0, -2.0), seed=k)
            for k in range(40)]
    srv = rt.Server(gpu_ctx, scs, max_active=6, mem_budget=64 << 20)
    try:
        alone = [srv.wait(srv.submit(**q, outputs=True), timeout_s=60.0) for q in reqs[:8]]
        tickets = []
        for k, q in enumerate(reqs):
            tickets.append(srv.submit(**q, outputs=True))
            if k % 5 == 4:
                srv.wait(tickets[k - 2], timeout_s=60.0)
                tickets[k - 2] = None
        res = [srv.wait(t, timeout_s=120.0) if t is not None else None for t in tickets]
        ticks, most = srv.counters()
    finally:
        srv.close()
    print(f"[serve] {ticks} ticks, largest tick {most} jobs", flush=True)
    for k in range(8):
        if res[k] is None:
            continue
        assert np.array_equal(alone[k]["id"], res[k]["id"]), f"request {k} ids"
        assert np.array_equal(alone[k]["occ"], res[k]["occ"]), f"request {k} occlusion"
    for s in scs:
        s.close()
