"""GPU parity: the HIP accelerators through the C ABI against the CPU oracle.

Bar: hit flag, triangle id and occlusion bit identical; t/u/v within 1e-5 relative (the
kernels use the oracle's FP32 operation order, so they are expected to match bit for bit,
which is also reported).  Full-size configurations are checked with size-independent
properties (every reported hit is a strict hit of that triangle at that t, occlusion
consistent with the closest hit, identical results on repeated calls).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _to_np(h):
    return dict(id=h.id.cpu().numpy(), t=h.t.cpu().numpy(), u=h.u.cpu().numpy(),
                v=h.v.cpu().numpy())


def _uv_from_b(b1, b2):
    b0 = (np.float32(1) - b1) - b2
    u = (b0 * np.float32(0) + b1 * np.float32(1)) + b2 * np.float32(1)
    v = (b0 * np.float32(0) + b1 * np.float32(0)) + b2 * np.float32(1)
    return u.astype(np.float32), v.astype(np.float32)


def _assert_closest(gpu, ref, label):
    print(f"[{label}] rays={gpu['id'].size} hits={int((ref['id'] >= 0).sum())}", flush=True)
    mism = np.nonzero(gpu["id"] != ref["id"])[0]
    assert mism.size == 0, f"{label}: {mism.size} id mismatches, first {mism[:8]}"
    hit = ref["id"] >= 0
    np.testing.assert_allclose(gpu["t"][hit], ref["t"][hit], rtol=RTOL, atol=0)
    u, v = _uv_from_b(ref["b1"], ref["b2"])
    np.testing.assert_allclose(gpu["u"][hit], u[hit], rtol=RTOL, atol=1e-7)
    np.testing.assert_allclose(gpu["v"][hit], v[hit], rtol=RTOL, atol=1e-7)
    assert np.all(np.isinf(gpu["t"][~hit]))
    exact = (np.array_equal(gpu["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32)))
    print(f"[{label}] t bit-exact: {exact}", flush=True)


def _dev(rt, r):
    return rt.DeviceRays.from_numpy(r)


# ---------------------------------------------------------------- naive
def test_naive_random1k(rt, scenes, orc, gpu_ctx):
    v9 = scenes.random_triangles(1024, seed=1)
    rays = scenes.pinhole_rays(128, 128)
    sc = rt.Scene(gpu_ctx, v9, desc="naive")
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    ref = orc.trace(v9, rays, orc.params("naive"), orc.CLOSEST_PRIMARY)
    _assert_closest(h, ref, "naive closest 128x128 x 1024")
    sh = scenes.shadow_rays_np(rays, h["id"], h["t"], (2, 2, -2))
    occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
    oref = orc.trace(v9, sh, orc.params("naive"), orc.OCCLUDED)["occ"]
    assert np.array_equal(occ, oref), f"{(occ != oref).sum()} occlusion mismatches"
    sc.close()


def test_naive_secondary_validity(rt, scenes, orc, gpu_ctx):
    v9 = scenes.random_triangles(700, seed=21)
    rays = scenes.random_rays(5000, seed=22, lo=-1.5, hi=1.5, maxt=np.inf)
    rays["valid"] = (np.arange(5000) % 4 != 0).astype(np.uint8)
    sc = rt.Scene(gpu_ctx, v9, desc="naive")
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_SECONDARY))
    ref = orc.trace(v9, rays, orc.params("naive"), orc.CLOSEST_SECONDARY)
    _assert_closest(h, ref, "naive secondary")
    sc.close()


# ---------------------------------------------------------------- rayhierarchy
@pytest.mark.parametrize("node", ["box_dir", "ia", "box_uv"])
def test_rayhierarchy_torus_window(rt, scenes, orc, gpu_ctx, node):
    """Torus scene (24,778 triangles), a 128x128 window of the 512x512 camera.  box_uv's cull
    tests the z range of the grown box's corner directions, which misses most hits of this
    head-on view below the camera: the reference's result, matched bit for bit."""
    v9 = scenes.torus_scene()
    full = scenes.pinhole_rays(512, 512)
    rays = scenes.subimage(full, 192, 256, 128, 128)
    desc = rt.default_desc("rayhierarchy", node=node)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    p = orc.params("rayhierarchy", node=node)
    ref = orc.trace(v9, rays, p, orc.CLOSEST_PRIMARY)
    _assert_closest(h, ref, f"rayhierarchy/{node} primary")
    sh = scenes.shadow_rays_np(rays, h["id"], h["t"], (2, 2, -2))
    occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
    oref = orc.trace(v9, sh, p, orc.OCCLUDED)["occ"]
    assert np.array_equal(occ, oref), f"{(occ != oref).sum()} occlusion mismatches"
    print(f"[rayhierarchy/{node}] shadow rays={int(sh['valid'].sum())} occluded={int(oref.sum())}")
    sc.close()


def test_rayhierarchy_random1k_full(rt, scenes, orc, gpu_ctx):
    v9 = scenes.random_triangles(1024, seed=1)
    rays = scenes.pinhole_rays(128, 128)
    sc = rt.Scene(gpu_ctx, v9, desc="rayhierarchy")
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    ref = orc.trace(v9, rays, orc.params("rayhierarchy"), orc.CLOSEST_PRIMARY)
    _assert_closest(h, ref, "rayhierarchy 128x128 x 1024")
    sc.close()


@pytest.mark.parametrize("W,H,cx,cy,height,node", [(37, 21, 4, 4, 3, "box_dir"),
                                                   (64, 48, 2, 4, 4, "box_dir"),
                                                   (40, 40, 8, 8, 2, "box_dir"),
                                                   (33, 17, 3, 5, 3, "box_dir"),
                                                   (37, 21, 4, 4, 3, "box_uv"),
                                                   (33, 17, 3, 5, 3, "box_uv")])
def test_rayhierarchy_shapes(rt, scenes, orc, gpu_ctx, W, H, cx, cy, height, node):
    """Ragged images, non-square leaves, partial tiles, other heights; box_uv with rays of
    d.x == 0 among the secondary and shadow rays (the validity-aware 5DB leaf build restarts
    its slope interval at them)."""
    v9 = scenes.random_triangles(600, seed=W * H, extent=1.0, size=0.2)
    rays = scenes.pinhole_rays(W, H, fov_deg=70)
    desc = rt.default_desc("rayhierarchy", chunk_x=cx, chunk_y=cy, height=height, node=node)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    p = orc.params("rayhierarchy", cx=cx, cy=cy, height=height, node=node)
    ref = orc.trace(v9, rays, p, orc.CLOSEST_PRIMARY)
    _assert_closest(h, ref, f"rayhierarchy {W}x{H} leaf {cx}x{cy} h{height}")
    # secondary rays with validity holes
    rays2 = scenes.random_rays(W * H, seed=5, lo=-1.2, hi=1.2, maxt=np.inf, width=W)
    rays2["valid"] = (np.random.default_rng(1).random(W * H) < 0.6).astype(np.uint8)
    if node == "box_uv":
        rays2["dx"][::5] = 0.0
    h2 = _to_np(sc.trace_closest(_dev(rt, rays2), rt.RT_CLOSEST_SECONDARY))
    ref2 = orc.trace(v9, rays2, p, orc.CLOSEST_SECONDARY)
    _assert_closest(h2, ref2, "rayhierarchy secondary")
    rays3 = dict(rays2, tmax=np.full(W * H, 1.0, np.float32))
    occ = sc.trace_occluded(_dev(rt, rays3)).cpu().numpy()
    oref = orc.trace(v9, rays3, p, orc.OCCLUDED)["occ"]
    assert np.array_equal(occ, oref)
    sc.close()


def test_rayhierarchy_repair_hit_at_tmin(rt, orc, gpu_ctx):
    """A triangle through the camera origin (t == tmin == 0) wins through the repair pass."""
    plane = [-1, -1, 0, 2, -1, 0, -1, 2, 0]
    far = [-1, -1, 5, 2, -1, 5, -1, 2, 5]
    v9 = np.array([far, plane], np.float32)
    n = 16
    z = np.zeros(n, np.float32)
    rays = dict(ox=z.copy(), oy=z.copy(), oz=z.copy(), dx=np.full(n, 0.01, np.float32),
                dy=np.full(n, 0.02, np.float32), dz=np.ones(n, np.float32), tmin=z.copy(),
                tmax=np.full(n, np.inf, np.float32), valid=None, width=4, height=4)
    sc = rt.Scene(gpu_ctx, v9, desc="rayhierarchy")
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    ref = orc.trace(v9, rays, orc.params("rayhierarchy"), orc.CLOSEST_PRIMARY)
    assert ref["stats"].repair_rays == n
    _assert_closest(h, ref, "repair at tmin")
    assert np.all(h["id"] == 1)
    sc.close()


# ---------------------------------------------------------------- ray-bvh
@pytest.mark.parametrize("node", ["box_dir", "ia", "box_uv"])
def test_raybvh_room_window(rt, scenes, orc, gpu_ctx, node):
    v9 = scenes.room_scene(nclusters=60)
    full = scenes.pinhole_rays(256, 256, origin=(0.0, 0.0, -9.0), fov_deg=50)
    rays = scenes.subimage(full, 64, 64, 96, 96)
    desc = rt.default_desc("ray-bvh", node=node)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    nodes, roots, order = sc.handoff
    p = orc.params("ray-bvh", node=node)
    h = _to_np(sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY))
    ref = orc.trace_raybvh(v9, nodes, roots, order, rays, p, orc.CLOSEST_PRIMARY)
    _assert_closest(h, ref, f"ray-bvh/{node} primary")
    sh = scenes.shadow_rays_np(rays, h["id"], h["t"], (1.0, 2.5, -8.0))
    occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
    oref = orc.trace_raybvh(v9, nodes, roots, order, sh, p, orc.OCCLUDED)["occ"]
    assert np.array_equal(occ, oref)
    rays2 = scenes.random_rays(96 * 96, seed=31, lo=-4, hi=4, maxt=np.inf, width=96)
    rays2["valid"] = (np.random.default_rng(2).random(96 * 96) < 0.8).astype(np.uint8)
    h2 = _to_np(sc.trace_closest(_dev(rt, rays2), rt.RT_CLOSEST_SECONDARY))
    ref2 = orc.trace_raybvh(v9, nodes, roots, order, rays2, p, orc.CLOSEST_SECONDARY)
    _assert_closest(h2, ref2, f"ray-bvh/{node} secondary")
    sc.close()


# ---------------------------------------------------------------- shading record, shadow rays
@pytest.mark.parametrize("accel", ["rayhierarchy", "naive"])
def test_shading_record_with_uvs(rt, scenes, orc, gpu_ctx, accel):
    """dpdu / dpdv / u / v of the hits.  Zero UV determinants take the coordinate-system
    fallback; determinants of +-0.001 take it only in the naive kernels (|det| < EPS,
    naive.cl:60-61), the others divide by them (computeDpTuTv.cl:57)."""
    v9 = scenes.random_triangles(500, seed=41)
    uv6 = np.random.default_rng(3).random((500, 6)).astype(np.float32)
    uv6[::7] = 0.0  # zero UV determinant -> coordinate-system fallback
    uv6[3::7] = [0, 0, 1, 0, 1, 0.001]    # determinant 0.001
    uv6[5::7] = [0, 0, 1, 0, 1, -0.001]   # determinant -0.001
    rays = scenes.pinhole_rays(64, 64)
    sc = rt.Scene(gpu_ctx, v9, uv6=uv6, desc=accel)
    hd = sc.trace_closest(_dev(rt, rays), rt.RT_CLOSEST_PRIMARY, shading=True)
    h = _to_np(hd)
    ref = orc.trace(v9, rays, orc.params(accel), orc.CLOSEST_PRIMARY)
    assert np.array_equal(h["id"], ref["id"])
    s = orc.shading(v9, uv6, rays, ref["id"], naive=accel == "naive")
    small = np.isin(ref["id"], np.r_[np.arange(3, 500, 7), np.arange(5, 500, 7)])
    assert small.sum() > 20, "hits on the small-determinant triangles"
    hit = ref["id"] >= 0
    np.testing.assert_allclose(h["u"][hit], s["u"][hit], rtol=RTOL, atol=1e-6)
    np.testing.assert_allclose(h["v"][hit], s["v"][hit], rtol=RTOL, atol=1e-6)
    du = hd.dpdu.cpu().numpy().reshape(-1, 3)
    dv = hd.dpdv.cpu().numpy().reshape(-1, 3)
    np.testing.assert_allclose(du[hit], s["dpdu"].reshape(-1, 3)[hit], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(dv[hit], s["dpdv"].reshape(-1, 3)[hit], rtol=1e-4, atol=1e-5)
    sc.close()


def test_shadow_ray_generation_matches_numpy(rt, scenes, gpu_ctx):
    v9 = scenes.random_triangles(1024, seed=1)
    rays = scenes.pinhole_rays(128, 128)
    sc = rt.Scene(gpu_ctx, v9, desc="naive")
    d = _dev(rt, rays)
    hd = sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY)
    s = rt.make_shadow_rays(d, hd, (2, 2, -2))
    h = _to_np(hd)
    ref = scenes.shadow_rays_np(rays, h["id"], h["t"], (2, 2, -2))
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"):
        assert np.array_equal(getattr(s, k).cpu().numpy(), ref[k]), k
    assert np.array_equal(s.valid.cpu().numpy(), ref["valid"])
    sc.close()


# ---------------------------------------------------------------- full-size properties
def test_torus_512_full_properties(rt, scenes, orc, gpu_ctx):
    """512x512 primary + shadow on the full torus scene: every hit is a strict hit at that t,
    repeated calls are identical, and shadow rays only exist for hits."""
    v9 = scenes.torus_scene()
    rays = scenes.pinhole_rays(512, 512)
    sc = rt.Scene(gpu_ctx, v9, desc="rayhierarchy")
    d = _dev(rt, rays)
    h1 = _to_np(sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY))
    h2 = _to_np(sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY))
    assert np.array_equal(h1["id"], h2["id"]) and np.array_equal(h1["t"], h2["t"])
    hit = np.nonzero(h1["id"] >= 0)[0]
    assert hit.size > 100000
    sample = np.random.default_rng(0).choice(hit, 3000, replace=False)
    for i in sample:
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        dd = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        ok, (t, _, _) = orc.mt("strict", o, dd, v9[h1["id"][i]])
        assert ok and t == h1["t"][i]
    sh = rt.make_shadow_rays(d, sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY), (2, 2, -2))
    occ = sc.trace_occluded(sh).cpu().numpy()
    assert np.all(occ[h1["id"] < 0] == 0)
    sc.close()


# ---------------------------------------------------------------- bounce rays (path integrator)
@pytest.mark.parametrize("accel", ["ray-bvh", "rayhierarchy"])
def test_bounce_rays_secondary_trace(rt, scenes, orc, gpu_ctx, accel):
    """Incoherent secondary rays: cosine-weighted bounces from the room window's primary hits,
    traced in secondary mode (validity from the bounce) and checked against the oracle on the
    rays read back from the device."""
    v9 = scenes.room_scene(nclusters=60)
    full = scenes.pinhole_rays(256, 256, origin=(0.0, 0.0, -9.0), fov_deg=50)
    rays = scenes.subimage(full, 64, 64, 96, 96)
    sc = rt.Scene(gpu_ctx, v9, desc=accel)
    d = _dev(rt, rays)
    h = sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY)
    b = rt.make_bounce_rays(sc, d, h, seed=7, depth=0)
    b2 = rt.make_bounce_rays(sc, d, h, seed=7, depth=0)
    assert all(torch_equal(x, y) for x, y in zip(
        (b.dx, b.dy, b.dz, b.valid), (b2.dx, b2.dy, b2.dz, b2.valid))), "not deterministic"
    br = {k: getattr(b, k).cpu().numpy() for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin",
                                                   "tmax", "valid")}
    br["width"], br["height"] = b.width, b.height
    ids = h.id.cpu().numpy()
    assert np.array_equal(br["valid"].astype(bool), ids >= 0)
    ok = ids >= 0
    dl = np.sqrt(br["dx"] ** 2 + br["dy"] ** 2 + br["dz"] ** 2)[ok]
    assert np.all(np.abs(dl - 1) < 1e-5)
    # direction on the side of the incoming ray
    tri = v9.reshape(-1, 9)[ids[ok]]
    n = np.cross(tri[:, 3:6] - tri[:, 0:3], tri[:, 6:9] - tri[:, 0:3])
    din = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)[ok]
    flip = np.where(np.sum(n * din, 1) > 0, -1.0, 1.0)[:, None]
    dout = np.stack([br["dx"], br["dy"], br["dz"]], 1)[ok]
    assert np.all(np.sum(n * flip * dout, 1) >= -1e-4)
    hb = _to_np(sc.trace_closest(b, rt.RT_CLOSEST_SECONDARY))
    p = orc.params(accel)
    if accel == "ray-bvh":
        nodes, roots, order = sc.handoff
        ref = orc.trace_raybvh(v9, nodes, roots, order, br, p, orc.CLOSEST_SECONDARY)
    else:
        ref = orc.trace(v9, br, p, orc.CLOSEST_SECONDARY)
    _assert_closest(hb, ref, f"{accel} bounce")
    sc.close()


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))
