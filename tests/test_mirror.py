"""The pbrt accelerator layer (pbrt-v2_amd/host: NaiveAccel, RayHieararchy, RayBVH,
MakeAccelerator) driven like pbrt's renderer through tests/cpp/libmirror_driver.so.

CPU tests check the factories' parameter handling (no device call is made before a Severe).
GPU tests check that the pbrt-side unpack reproduces the reference's conventions on top of the
oracle's hits: Intersection fields (triangle, primitiveId, dg.u/v, dpdu/dpdv), Ray::maxt
shortened to the hit distance, hit[] flags for bounce rays, '0'/'1' occlusion characters, and
the naive accelerator's two quirks (index 0 means "no hit", naive.cpp:298; a bounce ray whose
direction x is below -5 is skipped, naive.cl Intersection2).
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "cpp" / "libmirror_driver.so"
RTOL = 1e-5
SPLIT = 37  # triangles [0, SPLIT) belong to primitive 7, the rest to primitive 8


def _driver():
    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(DRIVER))
    lib.mirror_run.restype = ctypes.c_int
    lib.mirror_last_error.restype = ctypes.c_char_p
    return lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def run(accel, params, v9, rays, mode, hit=None, uv6=None, pre=None, calls=1):
    """Runs the accelerator on the ray dict; returns a dict of per-ray outputs, or raises
    RuntimeError with the Severe message."""
    lib = _driver()
    v9 = np.ascontiguousarray(v9, np.float32).reshape(-1)
    ntri = v9.size // 9
    n = rays["ox"].size
    r8 = np.stack([rays[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")],
                  axis=1).astype(np.float32)
    r8 = np.ascontiguousarray(r8)
    hit_io = (np.ones(n, np.uint8) if hit is None else hit.astype(np.uint8)).copy()
    tri = np.full(n, -2, np.int32)
    prim = np.full(n, -2, np.int32)
    maxt = np.zeros(n, np.float32)
    uv = np.zeros(2 * n, np.float32)
    dp = np.zeros(6 * n, np.float32)
    occ = np.zeros(n, np.uint8)
    uvk = None if uv6 is None else np.ascontiguousarray(uv6, np.float32).reshape(-1)
    xres, yres, spp = pre if pre else (0, 0, 1)
    print(f"[mirror] {accel} {params!r} mode={mode} rays={n} tris={ntri}", flush=True)
    rc = lib.mirror_run(accel.encode(), params.encode(), _p(v9), _p(uvk), ctypes.c_uint32(ntri),
                        ctypes.c_uint32(SPLIT), xres, yres, spp, _p(r8), ctypes.c_uint32(n), mode,
                        calls, _p(hit_io), _p(tri), _p(prim), _p(maxt), _p(uv), _p(dp), _p(occ))
    if rc != 0:
        raise RuntimeError(lib.mirror_last_error().decode())
    return dict(hit=hit_io.astype(bool), tri=tri, prim=prim, maxt=maxt, u=uv[0::2], v=uv[1::2],
                dpdu=dp.reshape(-1, 6)[:, :3], dpdv=dp.reshape(-1, 6)[:, 3:],
                occ=occ.view("S1").astype(str))


def expect_closest(out, ref_id, ref_t, rays, orc, v9, uv6, label):
    """Checks the unpacked Intersections against oracle ids / t and the oracle's shading."""
    hit = ref_id >= 0
    print(f"[{label}] hits={int(hit.sum())}", flush=True)
    assert np.array_equal(out["hit"], hit), f"{label}: hit flags differ"
    assert np.array_equal(out["tri"], np.where(hit, ref_id, -1)), f"{label}: triangle ids differ"
    assert np.array_equal(out["prim"], np.where(hit, np.where(ref_id < SPLIT, 7, 8), -1))
    np.testing.assert_allclose(out["maxt"][hit], ref_t[hit], rtol=RTOL, atol=0)
    assert np.array_equal(out["maxt"][~hit], rays["tmax"][~hit]), "maxt of a miss changed"
    s = orc.shading(v9, uv6, rays, np.where(hit, ref_id, -1).astype(np.int32))
    np.testing.assert_allclose(out["u"][hit], s["u"][hit], rtol=RTOL, atol=1e-6)
    np.testing.assert_allclose(out["v"][hit], s["v"][hit], rtol=RTOL, atol=1e-6)
    np.testing.assert_allclose(out["dpdu"][hit], s["dpdu"].reshape(-1, 3)[hit], rtol=1e-4,
                               atol=1e-5)
    np.testing.assert_allclose(out["dpdv"][hit], s["dpdv"].reshape(-1, 3)[hit], rtol=1e-4,
                               atol=1e-5)


def _closest_rays(r):
    """The layer traces closest-hit rays to infinity whatever their maxt (rayhierarchy.cpp:845),
    so the oracle gets the same rays with tmax = inf."""
    q = dict(r)
    q["tmax"] = np.full(r["ox"].size, np.inf, np.float32)
    return q


# ------------------------------------------------------------------------------------ CPU
@pytest.mark.parametrize("accel,params,msg", [
    ("kdtree-gpu", "", "unknown"),
    ("naive", "method=RayTri", "RayTri"),
    ("naive", "onGPU=false", "onGPU"),
    ("rayhierarchy", "node=box_xyz", "Unknown accelarator node type box_xyz"),
    ("ray-bvh", "splitmethod=middle", "middle"),
])
def test_factory_rejects(scenes, accel, params, msg):
    v9 = scenes.random_triangles(8, seed=1)
    rays = scenes.random_rays(4, seed=2)
    with pytest.raises(RuntimeError, match=msg):
        run(accel, params, v9, rays, 0)


@pytest.mark.parametrize("which,msg", [
    (0, "Called Intersect with unsupported aggregate!"),
    (1, "Called Intersect with unsupported aggregate!"),
    (2, "Called IntersectP with unsupported aggregate!"),
    (3, "Called MaxRaysPerCall with unsoppurted aggregate!"),
])
def test_scene_unsupported_aggregate(which, msg):
    """pbrt's Scene dispatches only to the three ray-space accelerators; any other aggregate
    is Severe() with the reference's messages (scene.cpp:55-120, scene.h:71-99; the misspelling
    is the reference's)."""
    lib = _driver()
    lib.mirror_scene_unsupported.restype = ctypes.c_int
    assert lib.mirror_scene_unsupported(which) == 1
    assert lib.mirror_last_error().decode() == msg


# ------------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("accel", ["naive", "rayhierarchy", "ray-bvh"])
def test_scene_dispatch(scenes, accel):
    """pbrt's Scene forwards to the accelerator (scene.cpp:70-120, scene.h:71-99):
    Scene::Intersect without a bounce is the primary call, Scene::Intersect(..., bounce) is the bounce call (naive gets
    bounce + 1, so Scene's bounce 0 is already a secondary call, naive.cpp:215), and
    Scene::IntersectP is IntersectP.  The accelerator calls themselves are pinned to the oracle
    by the tests above; here the Scene route must give the same outputs field for field."""
    v9 = scenes.random_triangles(800, seed=21)
    n = 64 * 64
    rays = scenes.random_rays(n, seed=22, lo=-1.5, hi=1.5, maxt=np.inf, width=64)
    rays["dx"][::11] = np.float32(-6.0)  # naive skips such bounce rays
    sel = np.random.default_rng(23).random(n) < 0.75
    params = "node=ia" if accel != "naive" else ""
    pre = (64, 64, 1)

    def same(a, b, label):
        for k in ("hit", "tri", "prim", "maxt", "u", "v", "dpdu", "dpdv"):
            assert np.array_equal(a[k], b[k]), f"{accel} {label}: {k} differs"

    same(run(accel, params, v9, rays, 3, pre=pre), run(accel, params, v9, rays, 0, pre=pre),
         "Scene::Intersect")
    same(run(accel, params, v9, rays, 4, hit=sel, pre=pre),
         run(accel, params, v9, rays, 1, hit=sel, pre=pre), "Scene::Intersect(bounce 0)")
    if accel == "naive":
        # the accelerator's own bounce 0 is the camera kernel
        same(run(accel, params, v9, rays, 6, pre=pre), run(accel, params, v9, rays, 0, pre=pre),
             "NaiveAccel::Intersect(bounce 0)")
    sh = dict(rays)
    sh["tmax"] = np.random.default_rng(24).uniform(0.5, 4.0, n).astype(np.float32)
    a = run(accel, params, v9, sh, 5, hit=sel, pre=pre)
    b = run(accel, params, v9, sh, 2, hit=sel, pre=pre)
    assert np.array_equal(a["occ"], b["occ"]), f"{accel}: Scene::IntersectP differs"
    assert np.count_nonzero(a["occ"] == "1") > 0


@pytest.mark.gpu
@pytest.mark.parametrize("accel,node", [("rayhierarchy", "box_dir"), ("rayhierarchy", "ia"),
                                        ("ray-bvh", "box_dir"), ("ray-bvh", "box_uv")])
def test_layer_primary_and_shadow(scenes, orc, accel, node):
    v9 = scenes.torus_scene() if accel == "rayhierarchy" else scenes.room_scene(nclusters=60)
    ntri = v9.shape[0]
    uv6 = np.random.default_rng(5).random((ntri, 6)).astype(np.float32)
    if accel == "rayhierarchy":
        full = scenes.pinhole_rays(512, 512)
        rays = scenes.subimage(full, 192, 256, 128, 128)
    else:
        full = scenes.pinhole_rays(256, 256, origin=(0.0, 0.0, -9.0), fov_deg=50)
        rays = scenes.subimage(full, 64, 64, 96, 96)
    w = 128 if accel == "rayhierarchy" else 96
    rays["width"], rays["height"] = w, w
    params = f"node={node}"
    out = run(accel, params, v9, rays, 0, uv6=uv6, pre=(w, w, 1), calls=2)
    p = orc.params(accel, node=node)
    cr = _closest_rays(rays)
    if accel == "ray-bvh":
        nodes, roots, order = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3)
        ref = orc.trace_raybvh(v9, nodes, roots, order, cr, p, orc.CLOSEST_PRIMARY)
    else:
        ref = orc.trace(v9, cr, p, orc.CLOSEST_PRIMARY)
    expect_closest(out, ref["id"], ref["t"], rays, orc, v9, uv6, f"{accel}/{node} primary")

    sh = scenes.shadow_rays_np(rays, ref["id"], ref["t"], (1.0, 2.5, -8.0))
    sh["width"], sh["height"] = w, w
    hit = sh["valid"].astype(bool)
    o = run(accel, params, v9, sh, 2, hit=hit, pre=(w, w, 1))
    if accel == "ray-bvh":
        oref = orc.trace_raybvh(v9, nodes, roots, order, sh, p, orc.OCCLUDED)["occ"]
    else:
        oref = orc.trace(v9, sh, p, orc.OCCLUDED)["occ"]
    assert set(np.unique(o["occ"])) <= {"0", "1"}
    assert np.array_equal(o["occ"] == "1", oref.astype(bool)), "occlusion characters differ"
    assert np.all(o["occ"][~hit] == "0")


@pytest.mark.gpu
def test_layer_sort_param(scenes, orc, rt):
    """RayHieararchy "bool sort": the triangles are traced in BVHAccel order
    (rayhierarchy.cpp:152-155), so among hits at equal t the smallest SORTED id wins; the
    Intersection still names the triangle by its input index.  The scene holds coincident
    duplicates (input ids i and n + i) so that the tie rule is exercised."""
    base = scenes.torus_scene()
    v9 = np.concatenate([base, base[:3000]], axis=0)
    order = rt.bvh_order(v9, 1)
    full = scenes.pinhole_rays(512, 512)
    rays = scenes.subimage(full, 192, 256, 128, 128)
    rays["width"], rays["height"] = 128, 128
    out = run("rayhierarchy", "sort=true", v9, rays, 0, pre=(128, 128, 1))
    p = orc.params("rayhierarchy")
    ref = orc.trace(v9[order], _closest_rays(rays), p, orc.CLOSEST_PRIMARY)
    ids = np.where(ref["id"] >= 0, order[np.maximum(ref["id"], 0)].astype(np.int32), -1)
    plain = orc.trace(v9, _closest_rays(rays), p, orc.CLOSEST_PRIMARY)["id"]
    print(f"[sort] rays whose winner differs from the unsorted order: "
          f"{int(np.count_nonzero(ids != plain))}", flush=True)
    expect_closest(out, ids, ref["t"], rays, orc, v9, None, "rayhierarchy sort")


@pytest.mark.gpu
def test_layer_bounce_rays_and_partial_batch(scenes, orc):
    """Bounce Intersect: hit[] selects the rays; a batch smaller than the preprocessed task
    image is traced as a padded image whose padding rays are invalid."""
    v9 = scenes.random_triangles(1500, seed=11)
    n = 96 * 96
    rays = scenes.random_rays(n, seed=12, lo=-1.5, hi=1.5, maxt=np.inf, width=96)
    sel = np.random.default_rng(4).random(n) < 0.7
    out = run("rayhierarchy", "", v9, rays, 1, hit=sel, pre=(96, 96, 1))
    r2 = dict(rays)
    r2["valid"] = sel.astype(np.uint8)
    ref = orc.trace(v9, r2, orc.params("rayhierarchy"), orc.CLOSEST_SECONDARY)
    expect_closest(out, ref["id"], ref["t"], rays, orc, v9, None, "bounce")
    assert not np.any(out["hit"][~sel])

    # partial batch: the last 200 rays of the task are missing
    m = n - 200
    part = {k: (v[:m] if isinstance(v, np.ndarray) else v) for k, v in rays.items()}
    part["width"], part["height"] = m, 1
    outp = run("rayhierarchy", "", v9, part, 0, pre=(96, 96, 1))
    pad = dict(rays)
    pad["valid"] = (np.arange(n) < m).astype(np.uint8)
    refp = orc.trace(v9, pad, orc.params("rayhierarchy"), orc.CLOSEST_SECONDARY)
    expect_closest(outp, refp["id"][:m], refp["t"][:m], part, orc, v9, None, "partial batch")


@pytest.mark.gpu
def test_layer_naive_quirks(scenes, orc):
    """naive: a hit on triangle 0 is reported as no hit; bounce rays with d.x < -5 are
    skipped even when hit[] selects them."""
    rnd = scenes.random_triangles(400, seed=13)
    big = np.array([[-0.4, -0.4, -2.0, 0.4, -0.4, -2.0, 0.0, 0.4, -2.0]], np.float32)
    v9 = np.concatenate([big, rnd.reshape(-1, 9)], axis=0)
    rays = scenes.pinhole_rays(64, 64)
    out = run("naive", "", v9, rays, 0)
    ref = orc.trace(v9, _closest_rays(rays), orc.params("naive"), orc.CLOSEST_PRIMARY)
    n0 = int((ref["id"] == 0).sum())
    assert n0 > 50, "the test scene must put triangle 0 in view"
    ids = np.where(ref["id"] == 0, -1, ref["id"])
    expect_closest(out, ids, ref["t"], rays, orc, v9, None, "naive primary")

    n = 4096
    br = scenes.random_rays(n, seed=14, lo=-1.5, hi=1.5, maxt=np.inf, width=64)
    br["dx"][::9] = np.float32(-6.0)  # not normalised on purpose
    sel = np.random.default_rng(6).random(n) < 0.8
    ob = run("naive", "", v9, br, 1, hit=sel)
    r2 = dict(br)
    r2["valid"] = (sel & ~(br["dx"] < -5)).astype(np.uint8)
    refb = orc.trace(v9, r2, orc.params("naive"), orc.CLOSEST_SECONDARY)
    ids = np.where(refb["id"] == 0, -1, refb["id"])
    expect_closest(ob, ids, refb["t"], br, orc, v9, None, "naive bounce")
    assert not np.any(ob["hit"][::9])


# --------------------------------------------------------------- breadth-first renderer loop
def _render(accel, v9, xres, yres, spp=1, max_depth=0, light=(2.0, 2.0, -2.0), params=""):
    lib = _driver()
    lib.mirror_render.restype = ctypes.c_int
    v9 = np.ascontiguousarray(v9, np.float32).reshape(-1)
    rgb = np.zeros(3 * xres * yres, np.float32)
    stats = np.zeros(6, np.int64)
    org = np.array([0.0, 0.0, -3.0], np.float32)
    lt = np.array(light, np.float32)
    rc = lib.mirror_render(accel.encode(), params.encode(), _p(v9), ctypes.c_uint32(v9.size // 9),
                           xres, yres, spp, _p(org), ctypes.c_float(60.0), _p(lt),
                           ctypes.c_float(10.0), ctypes.c_float(0.5), max_depth, _p(rgb),
                           _p(stats), None)
    if rc != 0:
        raise RuntimeError(lib.mirror_last_error().decode())
    return rgb.reshape(yres, xres, 3)[..., 0], stats


@pytest.mark.gpu
def test_breadth_first_directlighting_image(scenes, orc):
    """The breadth-first loop (two 512x512 tasks of a 1024x512 image at MaxRaysPerCall
    409600) renders the same direct-lighting image as the oracle's hits and occlusion on the
    same task images."""
    v9 = scenes.torus_scene()
    img, st = _render("rayhierarchy", v9, 1024, 512)
    print(f"[render] tasks {st[0]} ({st[1]}x{st[2]}), camera {st[3]}, shadow {st[4]}", flush=True)
    assert tuple(st[:3]) == (2, 2, 1)
    full = scenes.pinhole_rays(1024, 512)
    F = np.float32
    ref = np.zeros((512, 1024), np.float32)
    tri = v9.reshape(-1, 9)
    n = np.cross(tri[:, 3:6] - tri[:, 0:3], tri[:, 6:9] - tri[:, 0:3]).astype(F)
    nl = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    n = (n * (F(1) / nl)[:, None]).astype(F)
    for x0 in (0, 512):
        r = scenes.subimage(full, x0, 0, 512, 512)
        h = orc.trace(v9, r, orc.params("rayhierarchy"), orc.CLOSEST_PRIMARY)
        sh = scenes.shadow_rays_np(r, h["id"], h["t"], (2.0, 2.0, -2.0))
        occ = orc.trace(v9, sh, orc.params("rayhierarchy"), orc.OCCLUDED)["occ"]
        lit = (h["id"] >= 0) & (occ == 0)
        nn = n[np.maximum(h["id"], 0)]
        d = np.stack([r["dx"], r["dy"], r["dz"]], 1)
        flip = ((nn[:, 0] * d[:, 0] + nn[:, 1] * d[:, 1]) + nn[:, 2] * d[:, 2]) > 0
        nn = np.where(flip[:, None], -nn, nn)
        wi = np.stack([sh["dx"], sh["dy"], sh["dz"]], 1)
        c = np.maximum(F(0), (nn[:, 0] * wi[:, 0] + nn[:, 1] * wi[:, 1]) + nn[:, 2] * wi[:, 2])
        with np.errstate(divide="ignore", invalid="ignore"):  # misses: masked below
            L = ((F(0.5) * F(0.31830988618379067154)) * F(10.0)) * c / (sh["tmax"] * sh["tmax"])
        ref[:, x0:x0 + 512] = np.where(lit, L, F(0)).reshape(512, 512)
    assert np.count_nonzero(ref) > 10000
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_breadth_first_path_runs_all_accelerators(scenes):
    """Path tracing with two bounces issues Intersect(..., bounce) and IntersectP batches;
    the three accelerators agree pixel for pixel on a small scene where their candidate sets
    cover every hit (ia culls are conservative; naive tests every pair)."""
    v9 = scenes.random_triangles(600, seed=3)
    imgs = {}
    for accel, params in (("rayhierarchy", "node=ia"), ("ray-bvh", "node=ia")):
        imgs[accel], st = _render(accel, v9, 96, 64, spp=2, max_depth=2, params=params)
        assert st[5] > 0, "no bounce rays were traced"
    np.testing.assert_allclose(imgs["rayhierarchy"], imgs["ray-bvh"], rtol=1e-5, atol=1e-7)
