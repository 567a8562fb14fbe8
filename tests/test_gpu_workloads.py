"""GPU: the three large configurations at full size.

  uneven batch   64 scenes (1e3-3e5 triangles), 7.4 M camera + 7.5 M random shadow rays: the
                 batched form bit-identical to the 64 plain per-scene call pairs, sampled ray
                 blocks equal to the oracle, and the block path forced on a few scenes equal
                 to the default routing;
  room frame     the 80,012-triangle room with ray-bvh, 1024x1024 at 4 spp as the renderer's 16
                 task images of 1024x256 rays, path depth 3 (camera, three cosine-weighted
                 bounces, a shadow ray per hit at every depth): batched passes bit-identical to
                 the tasks traced one after another, sampled blocks of every pass equal to the
                 oracle over the whole scene, and at full size every reported hit a strict hit
                 of its triangle at its t, every occluded ray blocked by some triangle;
  serving        16 resident scenes (1e3-1e6 triangles), requests of 128^2-1024^2 at 1-4 spp:
                 each request identical alone and in the mix.
A 16x16 block of rays is a whole hierarchy subtree at the reference defaults, so the oracle's
trace of a block-aligned window is the reference's answer for those rays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _window(r, x0, y0, w, h):
    W = r["width"]
    out = {}
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax", "valid"):
        a = r.get(k)
        out[k] = None if a is None else a.reshape(-1, W)[y0:y0 + h, x0:x0 + w].reshape(-1).copy()
    out["width"], out["height"] = w, h
    return out


def _host(rays):
    """Host copy of a DeviceRays."""
    d = {k: getattr(rays, k).cpu().numpy() for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin",
                                                      "tmax")}
    d["valid"] = None if rays.valid is None else rays.valid.cpu().numpy()
    d["width"], d["height"] = rays.width, rays.height
    return d


def _blocks(g, r, count):
    """`count` random block-aligned 16x16 windows of ray image r: (window, flat ray indices)."""
    out = []
    for _ in range(count):
        bx = int(g.integers(0, r["width"] // 16)) * 16
        by = int(g.integers(0, r["height"] // 16)) * 16
        sel = (np.arange(by, by + 16)[:, None] * r["width"] + np.arange(bx, bx + 16)).ravel()
        out.append((_window(r, bx, by, 16, 16), sel))
    return out


def test_uneven_batch_full(rt, scenes, orc, gpu_ctx):
    import torch
    jobs = scenes.uneven_batch(64, seed0=100)
    print(f"[c4] {len(jobs)} scenes, {sum(j['v9'].shape[0] for j in jobs)} triangles", flush=True)
    scs = [rt.Scene(gpu_ctx, j["v9"], desc="rayhierarchy") for j in jobs]
    prims = [rt.DeviceRays.from_numpy(j["primary"]) for j in jobs]
    shads = [rt.DeviceRays.from_numpy(j["shadow"]) for j in jobs]
    plain_h = [scs[k].trace_closest(prims[k]) for k in range(64)]
    plain_o = [scs[k].trace_occluded(shads[k]) for k in range(64)]
    b = rt.Batch(gpu_ctx)
    bh = [b.add_closest(scs[k], prims[k]) for k in range(64)]
    bo = [b.add_occluded(scs[k], shads[k]) for k in range(64)]
    b.run()
    torch.cuda.synchronize()
    for k in range(64):
        for f in ("id", "t", "u", "v"):
            x, y = getattr(bh[k], f), getattr(plain_h[k], f)
            if x.dtype == torch.float32:
                x, y = x.view(torch.int32), y.view(torch.int32)
            assert int((x != y).sum()) == 0, f"scene {k}: batched {f} differs from plain"
        assert int((bo[k] != plain_o[k]).sum()) == 0, f"scene {k}: occlusion differs"
    print("[c4] batched == plain for all 128 jobs", flush=True)
    # oracle on sampled blocks of eight scenes
    g = np.random.default_rng(11)
    for k in range(0, 64, 8):
        ids = plain_h[k].id.cpu().numpy()
        occ = plain_o[k].cpu().numpy()
        p = orc.params("rayhierarchy")
        for win, sel in _blocks(g, jobs[k]["primary"], 1):
            ref = orc.trace(jobs[k]["v9"], win, p, orc.CLOSEST_PRIMARY)
            assert np.array_equal(ids[sel], ref["id"]), f"scene {k}: camera block vs oracle"
        for win, sel in _blocks(g, jobs[k]["shadow"], 1):
            ref = orc.trace(jobs[k]["v9"], win, p, orc.OCCLUDED)
            assert np.array_equal(occ[sel], ref["occ"]), f"scene {k}: shadow block vs oracle"
    # the block path forced on four scenes (their shadow jobs enumerate every candidate pair)
    for k in (0, 1, 2, 3):
        scs[k].set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, rt.RAY_PATH_OFF)
        scs[k].set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, rt.RAY_PATH_OFF)
        h2 = scs[k].trace_closest(prims[k])
        o2 = scs[k].trace_occluded(shads[k])
        assert int((h2.id != plain_h[k].id).sum()) == 0, f"scene {k}: block path ids"
        assert int((o2 != plain_o[k]).sum()) == 0, f"scene {k}: block path occlusion"
    b.close()
    for s in scs:
        s.close()


def test_room_frame_path3_full(rt, scenes, orc, gpu_ctx):
    import torch
    v9 = scenes.room_scene(seed=3)
    assert len(v9) == 80012
    sc = rt.Scene(gpu_ctx, v9, desc="ray-bvh")
    nodes, roots, order = sc.handoff
    light = (1.0, 2.5, -8.0)
    full = scenes.pinhole_rays(1024, 1024, origin=(0.0, 0.0, -9.0), fov_deg=50.0, spp=4, seed=3)
    tw, th = 1024 * 4 // 4, 1024 // 4
    tasks = [scenes.subimage(full, x * tw, y * th, tw, th) for y in range(4) for x in range(4)]
    del full
    prim = [rt.DeviceRays.from_numpy(r) for r in tasks]
    T = len(tasks)
    print(f"[c3] {T} tasks of {tw}x{th} rays, {len(v9)} triangles", flush=True)

    def run(batched):
        """Every pass of the frame; returns per depth (hits, shadow rays, occlusion, bounce)."""
        out = []
        cur = prim
        hits = None
        for depth in range(4):
            mode = rt.RT_CLOSEST_PRIMARY if depth == 0 else rt.RT_CLOSEST_SECONDARY
            if batched:
                bb = rt.Batch(gpu_ctx)
                hits = [bb.add_closest(sc, cur[k], mode) for k in range(T)]
                bb.run()
                bb.close()
            else:
                hits = [sc.trace_closest(cur[k], mode) for k in range(T)]
            sh = [rt.make_shadow_rays(cur[k], hits[k], light) for k in range(T)]
            if batched:
                bb = rt.Batch(gpu_ctx)
                occ = [bb.add_occluded(sc, sh[k]) for k in range(T)]
                bb.run()
                bb.close()
            else:
                occ = [sc.trace_occluded(sh[k]) for k in range(T)]
            nxt = None
            if depth < 3:
                nxt = [rt.make_bounce_rays(sc, cur[k], hits[k], seed=17 + k, depth=depth + 1)
                       for k in range(T)]
            out.append((cur, hits, sh, occ))
            cur = nxt
        torch.cuda.synchronize()
        return out

    A = run(False)
    B = run(True)
    for depth in range(4):
        for k in range(T):
            ha, hb = A[depth][1][k], B[depth][1][k]
            assert int((ha.id != hb.id).sum()) == 0, f"depth {depth} task {k}: ids"
            assert int((ha.t.view(torch.int32) != hb.t.view(torch.int32)).sum()) == 0
            assert int((A[depth][3][k] != B[depth][3][k]).sum()) == 0, f"depth {depth} occ"
    print("[c3] batched == plain for every pass", flush=True)
    g = np.random.default_rng(13)
    p = orc.params("ray-bvh")
    for depth in range(4):
        rays_d, hits, sh, occ = A[depth]
        mode = orc.CLOSEST_PRIMARY if depth == 0 else orc.CLOSEST_SECONDARY
        nchecked = 0
        for k in (0, 5, 10, 15):
            r = _host(rays_d[k])
            ids = hits[k].id.cpu().numpy()
            ts = hits[k].t.cpu().numpy()
            for win, sel in _blocks(g, r, 1):
                ref = orc.trace_raybvh(v9, nodes, roots, order, win, p, mode)
                assert np.array_equal(ids[sel], ref["id"]), f"depth {depth} task {k} vs oracle"
                nchecked += 1
            s_h = _host(sh[k])
            o = occ[k].cpu().numpy()
            for win, sel in _blocks(g, s_h, 1):
                ref = orc.trace_raybvh(v9, nodes, roots, order, win, p, orc.OCCLUDED)
                assert np.array_equal(o[sel], ref["occ"]), f"depth {depth} task {k} occ vs oracle"
            # full size: every reported hit of this task is a strict hit at its t (sampled)
            hit = np.nonzero(ids >= 0)[0]
            for i in g.choice(hit, size=min(200, hit.size), replace=False):
                ok, (t, _, _) = orc.mt("strict", (r["ox"][i], r["oy"][i], r["oz"][i]),
                                       (r["dx"][i], r["dy"][i], r["dz"][i]), v9[ids[i]])
                assert ok and np.float32(t) == ts[i], f"depth {depth} task {k} ray {i}"
            # every occluded ray is blocked by a triangle in (tmin, tmax): checked with the
            # exact CPU BVH on a sample (the reference predicate's hit set is a subset of it)
            occ_i = np.nonzero(o)[0]
            if occ_i.size:
                pick = g.choice(occ_i, size=min(200, occ_i.size), replace=False)
                sub = {f: (None if s_h[f] is None else s_h[f][pick]) for f in
                       ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax", "valid")}
                sub["width"], sub["height"] = len(pick), 1
                bvh = orc.CpuBvh(v9)
                assert int(bvh.occluded(sub).sum()) == len(pick), f"depth {depth} occlusion"
                bvh.close()
        print(f"[c3] depth {depth}: {nchecked} blocks per mode equal to the oracle", flush=True)
    sc.close()


def test_serving_16_scenes_alone_equals_mix(rt, scenes, gpu_ctx):
    scs = []
    for i in range(16):
        g = np.random.Generator(np.random.PCG64(200 + i))
        ntri = int(round(np.exp(g.uniform(np.log(1e3), np.log(1e6)))))
        scs.append(rt.Scene(gpu_ctx, scenes.cluster_scene(ntri, seed=200 + i),
                            desc="rayhierarchy"))
    g = np.random.default_rng(17)
    reqs = [dict(scene=int(g.integers(0, 16)), width=int(s), height=int(s),
                 spp=int(g.integers(1, 5)), origin=(0.0, 0.0, -12.0), fov_deg=50.0,
                 light=(0.0, 8.0, -8.0), seed=k)
            for k, s in enumerate(g.choice([128, 256, 512, 1024], size=24))]
    srv = rt.Server(gpu_ctx, scs, max_active=256)
    try:
        tickets = [srv.submit(**q, outputs=True) for q in reqs]
        mixed = [srv.wait(t, timeout_s=300.0) for t in tickets]
        _, most = srv.counters()
        alone = [srv.wait(srv.submit(**q, outputs=True), timeout_s=300.0) for q in reqs]
    finally:
        srv.close()
    print(f"[c5] 24 requests, largest tick {most} jobs", flush=True)
    assert most > 2
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a["id"], m["id"]), f"request {k}: ids alone vs mix"
        assert np.array_equal(a["t"].view(np.uint32), m["t"].view(np.uint32)), f"request {k}: t"
        assert np.array_equal(a["occ"], m["occ"]), f"request {k}: occlusion"
        assert a["shadow_rays"] == int((a["id"] >= 0).sum())
    for s in scs:
        s.close()
