"""GPU: the per-ray path against the block path and the oracle.

A block (top-level hierarchy node) whose root passes many triangle clusters is traced ray by
ray through the scene's search BVH, with the reference's exact cull chain evaluated only for
triangles that pass the strict leaf test.  The threshold parameters force either path for
every block (1: per-ray for all, RAY_PATH_OFF: block enumeration for all), so each case is
traced both ways and must agree bit for bit (id, t, u, v, occlusion), and agree with the
oracle where the oracle finishes in seconds.  Covered: rayhierarchy box_dir and ia, ray-bvh
box_dir and ia; primary (with the repair pass), secondary (validity; ray-bvh's first-primitive
rule) and occlusion; coherent camera rays and incoherent random rays; a 1024-ray block
hierarchy; the repair-at-tmin case; the default threshold (a mix of both paths in one call); a
closest hit whose provisional bound fails its chain (the lane traces the ray again).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(h):
    return dict(id=h.id.cpu().numpy(), t=h.t.cpu().numpy(), u=h.u.cpu().numpy(),
                v=h.v.cpu().numpy())


def _same(a, b, label):
    for k in ("id", "t", "u", "v"):
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        mism = int((x != y).sum())
        assert mism == 0, f"{label}: {mism} {k} mismatches"


def _paths(rt, sc, d, mode):
    """Results of one trace through the per-ray path, the block path and the default choice."""
    out = {}
    for name, thr in (("ray", 1), ("block", rt.RAY_PATH_OFF), ("default", 0)):
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, thr)
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, thr)
        if mode == 2:
            out[name] = sc.trace_occluded(d).cpu().numpy()
        else:
            out[name] = _np(sc.trace_closest(d, mode))
    sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, 0)
    sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, 0)
    return out


def _check_paths(out, mode, label):
    if mode == 2:
        for k in ("block", "default"):
            n = int((out["ray"] != out[k]).sum())
            assert n == 0, f"{label}: per-ray vs {k}: {n} occlusion mismatches"
    else:
        _same(out["ray"], out["block"], f"{label} per-ray vs block")
        _same(out["ray"], out["default"], f"{label} per-ray vs default")


def _oracle(orc, sc, accel, node, rays, mode, **kw):
    p = orc.params(accel, node=node, **kw)
    if accel == "ray-bvh":
        nodes, roots, order = sc.handoff
        return orc.trace_raybvh(sc.v9.reshape(-1, 9), nodes, roots, order, rays, p, mode)
    return orc.trace(sc.v9.reshape(-1, 9), rays, p, mode)


@pytest.mark.parametrize("accel", ["rayhierarchy", "ray-bvh"])
@pytest.mark.parametrize("node", ["box_dir", "ia", "box_uv"])
def test_per_ray_matches_block_and_oracle(rt, scenes, orc, gpu_ctx, accel, node):
    v9 = scenes.cluster_scene(6000, seed=11)
    sc = rt.Scene(gpu_ctx, v9, desc=rt.default_desc(accel, node=node))
    cam = scenes.pinhole_rays(64, 64, origin=(0.0, 0.0, -12.0), fov_deg=50.0)
    rnd = scenes.random_rays(64 * 64, seed=12, lo=-5.0, hi=5.0, maxt=np.inf, width=64)
    rnd["valid"] = (np.random.default_rng(3).random(64 * 64) < 0.8).astype(np.uint8)
    for mode, rays, what in ((0, cam, "camera"), (1, rnd, "random secondary")):
        label = f"{accel}/{node} {what}"
        print(f"[per-ray] {label}", flush=True)
        d = rt.DeviceRays.from_numpy(rays)
        out = _paths(rt, sc, d, mode)
        _check_paths(out, mode, label)
        ref = _oracle(orc, sc, accel, node, rays, mode)
        n = int((out["ray"]["id"] != ref["id"]).sum())
        assert n == 0, f"{label}: {n} id mismatches against the oracle"
        hit = ref["id"] >= 0
        assert np.array_equal(out["ray"]["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32))
        # shadow rays of the hits toward a light inside the cluster volume, and random
        # segments of length 10
        sh = scenes.shadow_rays_np(rays, ref["id"], ref["t"], (0.0, 1.0, -1.0))
        sh["width"], sh["height"] = rays["width"], rays["height"]
        seg = scenes.random_rays(64 * 64, seed=13, lo=-5.0, hi=5.0, maxt=10.0, width=64)
        for srays, sw in ((sh, "shadow"), (seg, "segments")):
            ds = rt.DeviceRays.from_numpy(srays)
            so = _paths(rt, sc, ds, 2)
            _check_paths(so, 2, f"{label} {sw}")
            oref = _oracle(orc, sc, accel, node, srays, orc.OCCLUDED)["occ"]
            n = int((so["ray"] != oref).sum())
            assert n == 0, f"{label} {sw}: {n} occlusion mismatches against the oracle"
    sc.close()


def test_per_ray_big_blocks_and_repair(rt, scenes, orc, gpu_ctx):
    """A height-4 hierarchy of 2x4 leaves (512-ray blocks, chunks 0-7 of a block) and the
    repair pass winning at t == tmin through the per-ray path."""
    v9 = scenes.random_triangles(1024, seed=1)
    desc = rt.default_desc("rayhierarchy", chunk_x=2, chunk_y=4, height=4)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    rays = scenes.pinhole_rays(128, 64)
    out = _paths(rt, sc, rt.DeviceRays.from_numpy(rays), 0)
    _check_paths(out, 0, "h4 camera")
    ref = orc.trace(v9, rays, orc.params("rayhierarchy", cx=2, height=4), orc.CLOSEST_PRIMARY)
    assert np.array_equal(out["ray"]["id"], ref["id"])
    sc.close()

    plane = [-1, -1, 0, 2, -1, 0, -1, 2, 0]
    far = [-1, -1, 5, 2, -1, 5, -1, 2, 5]
    v9 = np.array([far, plane], np.float32)
    n = 16
    z = np.zeros(n, np.float32)
    rays = dict(ox=z.copy(), oy=z.copy(), oz=z.copy(), dx=np.full(n, 0.01, np.float32),
                dy=np.full(n, 0.02, np.float32), dz=np.ones(n, np.float32), tmin=z.copy(),
                tmax=np.full(n, np.inf, np.float32), valid=None, width=4, height=4)
    sc = rt.Scene(gpu_ctx, v9, desc="rayhierarchy")
    out = _paths(rt, sc, rt.DeviceRays.from_numpy(rays), 0)
    _check_paths(out, 0, "repair at tmin")
    assert np.all(out["ray"]["id"] == 1)
    sc.close()


@pytest.mark.parametrize("node", ["box_dir", "box_uv"])
def test_per_ray_raybvh_first_primitive_rule(rt, scenes, orc, gpu_ctx, node):
    """ray-bvh with maxBVHPrim 4: 6DB secondary rays test only the first primitive of each BVH
    leaf (intersection6DB_BVH2.cl:88), 5DB ones every primitive (intersection5DB_BVH2.cl:84);
    primary and shadow rays every primitive."""
    v9 = scenes.cluster_scene(3000, seed=17)
    desc = rt.default_desc("ray-bvh", max_bvh_prim=4, node=node)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    rnd = scenes.random_rays(48 * 48, seed=18, lo=-5.0, hi=5.0, maxt=np.inf, width=48)
    for mode in (0, 1):
        out = _paths(rt, sc, rt.DeviceRays.from_numpy(rnd), mode)
        _check_paths(out, mode, f"maxBVHPrim 4 {node} mode {mode}")
        ref = _oracle(orc, sc, "ray-bvh", node, rnd, mode)
        assert np.array_equal(out["ray"]["id"], ref["id"]), f"mode {mode} vs oracle"
    sc.close()


@pytest.mark.parametrize("repair_run", [0, 1])
def test_per_ray_closest_retrace_after_failed_bound(rt, orc, gpu_ctx, repair_run):
    """The per-ray closest-hit kernel prunes with a queued candidate before its chain is
    evaluated.  Here the nearest strict hit of ray 0 (triangle A at t = 5) fails its chain:
    the root's origin box, grown by the other rays' origins at (100, 100, 0), sees A away from
    every ray direction (intersectsTri rejects it at the root), while the larger triangle B at
    t = 8 passes at both levels.  The pruning by A's t cut B away, so the lane must trace the
    ray again and find B, as the oracle and the block path do."""
    A = [(-1, -1, 5), (1, -1, 5), (0, 1, 5)]
    B = [(-10, -10, 8), (110, -10, 8), (-10, 110, 8)]
    v9 = np.array([np.ravel(A), np.ravel(B)], np.float32)
    o = np.array([(0, 0, 0), (100, 100, 0), (100, 100, 0), (100, 100, 0)], np.float32)
    d = np.array([(0, 0, 1)] * 4, np.float32)
    rays = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(4, np.float32),
                tmax=np.full(4, np.inf, np.float32), valid=None, width=2, height=2)
    desc = rt.default_desc("rayhierarchy", chunk_x=1, chunk_y=1, height=2, repair_run=repair_run)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    p = orc.params("rayhierarchy", cx=1, cy=1, height=2, repair_run=repair_run)
    ref = orc.trace(v9, rays, p, orc.CLOSEST_PRIMARY)
    assert list(ref["id"]) == [1, -1, -1, -1]
    out = _paths(rt, sc, rt.DeviceRays.from_numpy(rays), 0)
    _check_paths(out, 0, f"retrace repair {repair_run}")
    assert np.array_equal(out["ray"]["id"], ref["id"])
    assert np.array_equal(out["ray"]["t"].view(np.uint32), ref["t"].view(np.uint32))
    sc.close()
