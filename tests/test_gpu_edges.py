"""GPU edge cases through the C ABI, each against the oracle: empty scenes, single-ray and
all-invalid images, exactly coincident triangles (tie -> smallest id), degenerate triangles,
rays starting on a surface, repairRun 2, ray-bvh leaf sizes 2-4 with the first-primitive
rule of secondary rays, BVH heights 1 and 6, a scene of more than 60,000 triangles (two
BVH parts), hand-built subtrees at the 12-level depth limit of the ray-bvh traversal, and the
block shapes of the reference's own scene files (top-level blocks of 1,600-6,400 rays)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(rt, r):
    return rt.DeviceRays.from_numpy(r)


def _closest(sc, rt, rays, mode):
    h = sc.trace_closest(_dev(rt, rays), mode)
    return h.id.cpu().numpy(), h.t.cpu().numpy()


def _ref(orc, sc, v9, rays, p, mode):
    if sc.desc.kind == 2:
        nodes, roots, order = sc.handoff
        return orc.trace_raybvh(v9, nodes, roots, order, rays, p, mode)
    return orc.trace(v9, rays, p, mode)


@pytest.mark.parametrize("accel", ["naive", "rayhierarchy", "ray-bvh"])
def test_empty_scene_and_tiny_images(rt, scenes, gpu_ctx, accel):
    empty = np.zeros((0, 9), np.float32)
    sc = rt.Scene(gpu_ctx, empty, desc=accel)
    for W, H in ((1, 1), (3, 1), (1, 5), (17, 3)):
        rays = scenes.pinhole_rays(W, H)
        ids, ts = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
        assert np.all(ids == -1) and np.all(np.isinf(ts))
        occ = sc.trace_occluded(_dev(rt, dict(rays, tmax=np.ones(W * H, np.float32)))).cpu()
        assert int(occ.sum()) == 0
    sc.close()


@pytest.mark.parametrize("accel", ["naive", "rayhierarchy", "ray-bvh"])
def test_single_ray_and_all_invalid(rt, scenes, orc, gpu_ctx, accel):
    v9 = scenes.random_triangles(300, seed=4)
    sc = rt.Scene(gpu_ctx, v9, desc=accel)
    p = orc.params(accel)
    one = scenes.pinhole_rays(1, 1, fov_deg=5)
    ids, _ = _closest(sc, rt, one, rt.RT_CLOSEST_PRIMARY)
    assert np.array_equal(ids, _ref(orc, sc, v9, one, p, orc.CLOSEST_PRIMARY)["id"])
    rr = scenes.random_rays(64 * 16, seed=9, lo=-1, hi=1, maxt=np.inf, width=64)
    rr["valid"] = np.zeros(64 * 16, np.uint8)
    ids, ts = _closest(sc, rt, rr, rt.RT_CLOSEST_SECONDARY)
    assert np.all(ids == -1) and np.all(np.isinf(ts))
    occ = sc.trace_occluded(_dev(rt, dict(rr, tmax=np.full(64 * 16, 5, np.float32)))).cpu()
    assert int(occ.sum()) == 0
    sc.close()


@pytest.mark.parametrize("accel", ["rayhierarchy", "ray-bvh", "naive"])
def test_coincident_triangles_tie_to_smallest_id(rt, scenes, orc, gpu_ctx, accel):
    """Three copies of the same triangle at ids 5, 9 and 200: every hit reports id 5."""
    v9 = scenes.random_triangles(256, seed=12).copy()
    big = np.array([-0.6, -0.6, -1.5, 0.7, -0.5, -1.5, 0.0, 0.7, -1.5], np.float32)
    for i in (5, 9, 200):
        v9[i] = big
    rays = scenes.pinhole_rays(64, 64)
    sc = rt.Scene(gpu_ctx, v9, desc=accel)
    ids, _ = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
    ref = _ref(orc, sc, v9, rays, orc.params(accel), orc.CLOSEST_PRIMARY)
    assert np.array_equal(ids, ref["id"])
    assert int((ids == 5).sum()) > 200 and not np.any((ids == 9) | (ids == 200))
    sc.close()


@pytest.mark.parametrize("accel", ["rayhierarchy", "ray-bvh"])
def test_degenerate_triangles_and_rays_on_surfaces(rt, scenes, orc, gpu_ctx, accel):
    """Zero-area triangles (a point, a segment) never hit; secondary rays leaving a surface
    at tmin = 0 exercise t == tmin."""
    v9 = scenes.random_triangles(400, seed=21).copy()
    v9[7] = np.array([0.1, 0.1, 0.1] * 3, np.float32)
    v9[8] = np.array([0, 0, 0, 0.5, 0.5, 0.5, 1, 1, 1], np.float32)
    rays = scenes.pinhole_rays(48, 48)
    sc = rt.Scene(gpu_ctx, v9, desc=accel)
    p = orc.params(accel)
    ids, ts = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
    ref = _ref(orc, sc, v9, rays, p, orc.CLOSEST_PRIMARY)
    assert np.array_equal(ids, ref["id"]) and not np.any((ids == 7) | (ids == 8))
    hit = ids >= 0
    t = np.where(hit, ts, 0).astype(np.float32)
    on = dict(ox=rays["ox"] + rays["dx"] * t, oy=rays["oy"] + rays["dy"] * t,
              oz=rays["oz"] + rays["dz"] * t, dx=rays["dx"], dy=rays["dy"], dz=rays["dz"],
              tmin=np.zeros(rays["ox"].size, np.float32),
              tmax=np.full(rays["ox"].size, np.inf, np.float32), valid=hit.astype(np.uint8),
              width=48, height=48)
    on = {k: (v.astype(np.float32) if isinstance(v, np.ndarray) and v.dtype != np.uint8 else v)
          for k, v in on.items()}
    ids2, _ = _closest(sc, rt, on, rt.RT_CLOSEST_SECONDARY)
    assert np.array_equal(ids2, _ref(orc, sc, v9, on, p, orc.CLOSEST_SECONDARY)["id"])
    sc.close()


def test_repair_run_two(rt, scenes, orc, gpu_ctx):
    v9 = scenes.torus_scene()
    rays = scenes.subimage(scenes.pinhole_rays(512, 512), 200, 200, 64, 64)
    desc = rt.default_desc("rayhierarchy", repair_run=2)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    ids, ts = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
    ref = orc.trace(v9, rays, orc.params("rayhierarchy", repair_run=2), orc.CLOSEST_PRIMARY)
    assert np.array_equal(ids, ref["id"])
    np.testing.assert_allclose(ts[ids >= 0], ref["t"][ids >= 0], rtol=1e-5)
    sc.close()


@pytest.mark.parametrize("max_prim,bvh_height", [(2, 3), (4, 3), (1, 1), (1, 6)])
def test_raybvh_leaf_sizes_and_heights(rt, scenes, orc, gpu_ctx, max_prim, bvh_height):
    """maxBVHPrim 2-4 (secondary rays test only the first primitive of a BVH leaf,
    intersection6DB_BVH2.cl) and other BVHheight frontiers."""
    v9 = scenes.room_scene(nclusters=40)
    full = scenes.pinhole_rays(256, 256, origin=(0.0, 0.0, -9.0), fov_deg=50)
    rays = scenes.subimage(full, 96, 96, 64, 64)
    desc = rt.default_desc("ray-bvh", max_bvh_prim=max_prim, bvh_height=bvh_height)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    p = orc.params("ray-bvh")
    ids, _ = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
    assert np.array_equal(ids, _ref(orc, sc, v9, rays, p, orc.CLOSEST_PRIMARY)["id"])
    rr = scenes.random_rays(64 * 64, seed=31, lo=-4, hi=4, maxt=np.inf, width=64)
    ids2, _ = _closest(sc, rt, rr, rt.RT_CLOSEST_SECONDARY)
    assert np.array_equal(ids2, _ref(orc, sc, v9, rr, p, orc.CLOSEST_SECONDARY)["id"])
    sh = scenes.shadow_rays_np(rays, ids, _closest(sc, rt, rays, 0)[1], (1.0, 2.5, -8.0))
    occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
    assert np.array_equal(occ, _ref(orc, sc, v9, sh, p, orc.OCCLUDED)["occ"])
    sc.close()


def test_raybvh_two_parts(rt, scenes, orc, gpu_ctx):
    """More than 60,000 triangles: the handoff holds two independently built BVH parts."""
    v9 = scenes.room_scene(nclusters=320)  # 64,012 triangles
    assert v9.shape[0] > 60000
    full = scenes.pinhole_rays(256, 256, origin=(0.0, 0.0, -9.0), fov_deg=50)
    rays = scenes.subimage(full, 112, 112, 32, 32)
    sc = rt.Scene(gpu_ctx, v9, desc="ray-bvh")
    ids, _ = _closest(sc, rt, rays, rt.RT_CLOSEST_PRIMARY)
    assert np.array_equal(ids, _ref(orc, sc, v9, rays, orc.params("ray-bvh"),
                                    orc.CLOSEST_PRIMARY)["id"])
    sc.close()



def _comb_bvh(rt, v9, order, nleaves):
    """Two comb-shaped subtrees of nleaves leaves each: interior node 2k has the leaf 2k+1 and
    the node 2k+2 as children, so the last two leaves sit nleaves-1 levels below their root.
    One triangle per leaf, BVH position p holding triangle order[p]; each box is the exact
    union of the triangle bounds below it."""
    tri = v9.reshape(-1, 9)
    lo = np.minimum(np.minimum(tri[:, 0:3], tri[:, 3:6]), tri[:, 6:9])
    hi = np.maximum(np.maximum(tri[:, 0:3], tri[:, 3:6]), tri[:, 6:9])
    nodes = np.zeros(2 * (2 * nleaves - 1), rt.GPU_NODE_DTYPE)
    roots = []
    for c in range(2):
        nb = c * (2 * nleaves - 1)
        roots.append(nb)
        box = {}
        for k in range(nleaves):  # leaf k at 2k+1, the last one at 2(nleaves-1)
            idx = nb + (2 * k + 1 if k < nleaves - 1 else 2 * k)
            pos = c * nleaves + k
            box[idx] = (lo[order[pos]], hi[order[pos]])
            nodes[idx] = (*box[idx][0], *box[idx][1], pos, 1)
        for k in range(nleaves - 2, -1, -1):
            a, b = box[nb + 2 * k + 1], box[nb + 2 * k + 2]
            box[nb + 2 * k] = (np.minimum(a[0], b[0]), np.maximum(a[1], b[1]))
            nodes[nb + 2 * k] = (*box[nb + 2 * k][0], *box[nb + 2 * k][1], nb + 2 * k + 1, 0)
    return nodes, np.array(roots, np.uint32)


def _comb_scene(nleaves, seed):
    g = np.random.default_rng(seed)
    n = 2 * nleaves
    centre = np.concatenate([g.uniform(-0.8, 0.8, (n, 2)), g.uniform(1.0, 3.0, (n, 1))], 1)
    v9 = (centre[:, None, :] + g.uniform(-0.6, 0.6, (n, 3, 3))).astype(np.float32)
    return v9.reshape(-1, 9), g.permutation(n).astype(np.uint32)


@pytest.mark.parametrize("node", ["box_dir", "ia", "box_uv"])
def test_raybvh_depth_limit_subtrees(rt, scenes, orc, gpu_ctx, node):
    """ray-bvh over frontier subtrees 12 levels deep, the deepest rt_scene_set_bvh accepts:
    the block traversal's bounded stack and pair queue hold, so the per-ray, block and
    default paths agree with the oracle bit for bit and the context's status word
    (rt_ctx_check) stays 0.  Subtrees 13 levels deep are rejected."""
    v9, order = _comb_scene(13, seed=71)
    nodes, roots = _comb_bvh(rt, v9, order, 13)
    gpu_ctx.check()  # clear what earlier traces recorded
    sc = rt.Scene(gpu_ctx, v9, desc=rt.default_desc("ray-bvh", node=node), build_bvh=False)
    sc.set_bvh(nodes, roots, order)
    p = orc.params("ray-bvh", node=node)
    cam = scenes.pinhole_rays(64, 64, origin=(0.0, 0.0, -3.0), fov_deg=24)
    # incoherent secondary rays: camera directions in random order from jittered origins
    g = np.random.default_rng(72)
    perm = g.permutation(64 * 64)
    rnd = {k: (v[perm].copy() if isinstance(v, np.ndarray) else v) for k, v in cam.items()}
    for k in ("ox", "oy", "oz"):
        rnd[k] = (rnd[k] + g.uniform(-0.3, 0.3, 64 * 64)).astype(np.float32)
    rnd["valid"] = (g.random(64 * 64) < 0.9).astype(np.uint8)
    for thr in (1, rt.RAY_PATH_OFF, 0):  # per-ray, block, the default choice
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, thr)
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, thr)
        for mode, rays in ((rt.RT_CLOSEST_PRIMARY, cam), (rt.RT_CLOSEST_SECONDARY, rnd)):
            ids, ts = _closest(sc, rt, rays, mode)
            ref = orc.trace_raybvh(v9, nodes, roots, order, rays, p, mode)
            assert np.array_equal(ids, ref["id"]), (thr, mode)
            assert np.array_equal(ts.view(np.uint32), ref["t"].view(np.uint32)), (thr, mode)
            assert (ids >= 0).sum() > 300, (thr, mode)
        ids, ts = _closest(sc, rt, cam, rt.RT_CLOSEST_PRIMARY)
        sh = scenes.shadow_rays_np(cam, ids, ts, (0.3, 2.0, -1.0))
        occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
        ref = orc.trace_raybvh(v9, nodes, roots, order, sh, p, orc.OCCLUDED)["occ"]
        assert np.array_equal(occ, ref), thr
        assert 0 < ref.sum() < (ids >= 0).sum()
        assert gpu_ctx.check() == 0, thr
    sc.close()
    v9d, orderd = _comb_scene(14, seed=73)
    deep = rt.Scene(gpu_ctx, v9d, desc=rt.default_desc("ray-bvh", node=node), build_bvh=False)
    with pytest.raises(rt.RtError) as e:
        deep.set_bvh(*_comb_bvh(rt, v9d, orderd, 14), orderd)
    assert e.value.status == 4  # RT_ERR_UNSUPPORTED
    deep.close()


@pytest.mark.parametrize("accel,node,W,H,cx,cy,height", [
    # prt-teapot.pbrt / sibenik.pbrt: 100x4 leaves, height 4, at 128x128 (blocks of 200x8)
    ("rayhierarchy", "ia", 128, 128, 100, 4, 4),
    ("rayhierarchy", "box_uv", 128, 128, 100, 4, 4),
    # conference.pbrt: 100x4 leaves, height 3 (three levels here: blocks of 400x16)
    ("rayhierarchy", "box_dir", 320, 64, 100, 4, 3),
    # the ray-bvh lines of the same files: 32x2 leaves, height 5 (four levels: 256x16)
    ("ray-bvh", "box_dir", 256, 64, 32, 2, 5),
    # prt-teapot.pbrt's ray-bvh line: 2x2 leaves, height 5, BVHheight 5 (five levels)
    ("ray-bvh", "ia", 128, 128, 2, 2, 5),
    # six levels of 2x2 leaves (blocks of 64x64 rays)
    ("rayhierarchy", "box_dir", 128, 128, 2, 2, 6),
])
def test_reference_scene_block_shapes(rt, scenes, orc, gpu_ctx, accel, node, W, H, cx, cy,
                                      height):
    """Top-level blocks of more than 1,024 rays or more than four levels, as the reference's
    scene files configure them: they are built in global memory where deeper than four levels,
    take the per-ray path whatever the routing thresholds say, and match the oracle in every
    mode (primary with repair, secondary with validity, occlusion), alone and batched with a
    job of small blocks."""
    v9 = scenes.room_scene(nclusters=40)
    extra = {"bvh_height": 5} if (accel == "ray-bvh" and cx == 2) else {}
    desc = rt.default_desc(accel, node=node, chunk_x=cx, chunk_y=cy, height=height, **extra)
    sc = rt.Scene(gpu_ctx, v9, desc=desc)
    p = orc.params(accel, node=node, cx=cx, cy=cy, height=height)
    cam = scenes.pinhole_rays(W, H, origin=(0.0, 0.0, -9.0), fov_deg=50)
    rnd = scenes.random_rays(W * H, seed=31, lo=-4, hi=4, maxt=np.inf, width=W)
    rnd["valid"] = (np.arange(W * H) % 7 != 0).astype(np.uint8)
    refs = {mode: _ref(orc, sc, v9, rays, p, mode)
            for mode, rays in ((rt.RT_CLOSEST_PRIMARY, cam), (rt.RT_CLOSEST_SECONDARY, rnd))}
    sh = scenes.shadow_rays_np(cam, refs[0]["id"], refs[0]["t"], (1.0, 2.5, -8.0))
    ref_occ = _ref(orc, sc, v9, sh, p, orc.OCCLUDED)["occ"]
    assert ref_occ.sum() > 500
    for thr in (0, rt.RAY_PATH_OFF):
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, thr)
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, thr)
        for mode, rays in ((rt.RT_CLOSEST_PRIMARY, cam), (rt.RT_CLOSEST_SECONDARY, rnd)):
            ids, ts = _closest(sc, rt, rays, mode)
            assert np.array_equal(ids, refs[mode]["id"]), (thr, mode)
            assert np.array_equal(ts.view(np.uint32), refs[mode]["t"].view(np.uint32)), (thr, mode)
        occ = sc.trace_occluded(_dev(rt, sh)).cpu().numpy()
        assert np.array_equal(occ, ref_occ), thr
    sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, 0)
    sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, 0)
    # one batch: this job and a job of 4x4 leaves (256-ray blocks) in the same launch group
    small = rt.Scene(gpu_ctx, v9, desc=rt.default_desc(accel, node=node))
    b = rt.Batch(gpu_ctx)
    dcam = _dev(rt, cam)
    h_big = b.add_closest(sc, dcam, rt.RT_CLOSEST_PRIMARY)
    h_small = b.add_closest(small, dcam, rt.RT_CLOSEST_PRIMARY)
    b.run()
    import torch
    torch.cuda.synchronize()
    ref_small = _ref(orc, small, v9, cam, orc.params(accel, node=node), orc.CLOSEST_PRIMARY)
    assert np.array_equal(h_big.id.cpu().numpy(), refs[0]["id"])
    assert np.array_equal(h_small.id.cpu().numpy(), ref_small["id"])
    b.close()
    small.close()
    sc.close()
