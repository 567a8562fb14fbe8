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


def expect_closest(out, ref_id, ref_t, rays, orc, v9, uv6, label, naive=False):
    """Checks the unpacked Intersections against oracle ids / t and the oracle's shading
    (naive: the naive kernels' |det| < EPS fallback)."""
    hit = ref_id >= 0
    print(f"[{label}] hits={int(hit.sum())}", flush=True)
    assert np.array_equal(out["hit"], hit), f"{label}: hit flags differ"
    assert np.array_equal(out["tri"], np.where(hit, ref_id, -1)), f"{label}: triangle ids differ"
    assert np.array_equal(out["prim"], np.where(hit, np.where(ref_id < SPLIT, 7, 8), -1))
    np.testing.assert_allclose(out["maxt"][hit], ref_t[hit], rtol=RTOL, atol=0)
    assert np.array_equal(out["maxt"][~hit], rays["tmax"][~hit]), "maxt of a miss changed"
    s = orc.shading(v9, uv6, rays, np.where(hit, ref_id, -1).astype(np.int32), naive=naive)
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
@pytest.mark.parametrize("accel,node,split", [("rayhierarchy", "box_dir", "sah"),
                                              ("rayhierarchy", "ia", "sah"),
                                              ("ray-bvh", "box_dir", "sah"),
                                              ("ray-bvh", "box_uv", "sah"),
                                              ("ray-bvh", "box_dir", "middle"),
                                              ("ray-bvh", "box_dir", "equal")])
def test_layer_primary_and_shadow(scenes, orc, accel, node, split):
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
    params = f"node={node};splitmethod={split}"
    out = run(accel, params, v9, rays, 0, uv6=uv6, pre=(w, w, 1), calls=2)
    p = orc.params(accel, node=node)
    cr = _closest_rays(rays)
    if accel == "ray-bvh":
        nodes, roots, order = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3, split=split)
        ref = orc.trace_raybvh(v9, nodes, roots, order, cr, p, orc.CLOSEST_PRIMARY)
    else:
        ref = orc.trace(v9, cr, p, orc.CLOSEST_PRIMARY)
    expect_closest(out, ref["id"], ref["t"], rays, orc, v9, uv6,
                   f"{accel}/{node}/{split} primary")

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
@pytest.mark.parametrize("split", ["sah", "middle", "equal"])
def test_layer_sort_param(scenes, orc, rt, split):
    """RayHieararchy "bool sort": the triangles are traced in BVHAccel order
    (rayhierarchy.cpp:152-155) built with the "splitmethod" parameter, so among hits at equal t
    the smallest SORTED id wins; the Intersection still names the triangle by its input index.
    The scene holds coincident duplicates (input ids i and n + i) so that the tie rule is
    exercised."""
    base = scenes.torus_scene()
    v9 = np.concatenate([base, base[:3000]], axis=0)
    order = rt.bvh_order(v9, 1, split=split)
    full = scenes.pinhole_rays(512, 512)
    rays = scenes.subimage(full, 192, 256, 128, 128)
    rays["width"], rays["height"] = 128, 128
    out = run("rayhierarchy", f"sort=true;splitmethod={split}", v9, rays, 0, pre=(128, 128, 1))
    p = orc.params("rayhierarchy")
    ref = orc.trace(v9[order], _closest_rays(rays), p, orc.CLOSEST_PRIMARY)
    ids = np.where(ref["id"] >= 0, order[np.maximum(ref["id"], 0)].astype(np.int32), -1)
    plain = orc.trace(v9, _closest_rays(rays), p, orc.CLOSEST_PRIMARY)["id"]
    print(f"[sort] rays whose winner differs from the unsorted order: "
          f"{int(np.count_nonzero(ids != plain))}", flush=True)
    expect_closest(out, ids, ref["t"], rays, orc, v9, None, f"rayhierarchy sort {split}")


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
    expect_closest(out, ids, ref["t"], rays, orc, v9, None, "naive primary", naive=True)

    n = 4096
    br = scenes.random_rays(n, seed=14, lo=-1.5, hi=1.5, maxt=np.inf, width=64)
    br["dx"][::9] = np.float32(-6.0)  # not normalised on purpose
    sel = np.random.default_rng(6).random(n) < 0.8
    ob = run("naive", "", v9, br, 1, hit=sel)
    r2 = dict(br)
    r2["valid"] = (sel & ~(br["dx"] < -5)).astype(np.uint8)
    refb = orc.trace(v9, r2, orc.params("naive"), orc.CLOSEST_SECONDARY)
    ids = np.where(refb["id"] == 0, -1, refb["id"])
    expect_closest(ob, ids, refb["t"], br, orc, v9, None, "naive bounce", naive=True)
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


def _mix32(h):
    h = h.astype(np.uint32)
    h ^= h >> np.uint32(16)
    h = (h * np.uint32(0x7feb352d)).astype(np.uint32)
    h ^= h >> np.uint32(15)
    h = (h * np.uint32(0x846ca68b)).astype(np.uint32)
    h ^= h >> np.uint32(16)
    return h


def _libm(name, argtype):
    fn = getattr(ctypes.CDLL("libm.so.6"), name)
    fn.restype = fn.argtypes = None
    fn.restype = argtype
    fn.argtypes = [argtype]
    return fn


def _oracle_path_image(orc, v9, accel, params, xres, yres, spp, max_depth, light=(2.0, 2.0, -2.0),
                       intensity=10.0, albedo=0.5, seed=1):
    """The breadth-first path loop restated over the oracle's traces (one task: the whole image
    in sampler order): camera rays through (x + (k + 0.5) / spp, y + 0.5), a shadow ray per hit
    toward the point light (mint 1e-3 t, maxt the distance), Lambertian direct light, and
    cosine-weighted bounces from the mix32 hash of (seed, ray, depth, task), traced as bounce
    rays of the hits.  FP32 in the renderer's operation order; cosf / sinf / tan from libm,
    as the renderer calls them."""
    F = np.float32
    cosf, sinf, tan = _libm("cosf", ctypes.c_float), _libm("sinf", ctypes.c_float), \
        _libm("tan", ctypes.c_double)
    tri = v9.reshape(-1, 3, 3)
    e1, e2 = (tri[:, 1] - tri[:, 0]).astype(F), (tri[:, 2] - tri[:, 0]).astype(F)
    nrm = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                    e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1).astype(F)
    ln = np.sqrt((nrm[:, 0] * nrm[:, 0] + nrm[:, 1] * nrm[:, 1]) + nrm[:, 2] * nrm[:, 2])
    nrm = np.where(ln[:, None] == 0, nrm, nrm * (F(1) / np.where(ln == 0, F(1), ln))[:, None]).astype(F)
    tan_half = F(tan(60.0 * (3.14159265358979323846 / 180.0) / 2.0))
    aspect = F(xres / yres)
    ys, xs, ks = np.meshgrid(np.arange(yres), np.arange(xres), np.arange(spp), indexing="ij")
    fx = xs.astype(F) + ((ks.astype(F) + F(0.5)) / F(spp))
    fy = ys.astype(F) + F(0.5)
    px = (((fx / F(xres)) * F(2) - F(1)) * tan_half * aspect).reshape(-1)
    py = ((F(1) - (fy / F(yres)) * F(2)) * tan_half).reshape(-1)
    pz = np.ones_like(px)
    inv = F(1) / np.sqrt((px * px + py * py) + pz * pz)
    n = px.size
    o = np.tile(np.array([0.0, 0.0, -3.0], F), (n, 1))
    d = np.stack([px * inv, py * inv, pz * inv], 1).astype(F)
    W, H = xres * spp, yres
    kind = {"naive": 0, "rayhierarchy": 1, "ray-bvh": 2}[accel]
    p = orc.params(accel, **params)
    handoff = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3) if kind == 2 else None

    def trace(o, d, mint, maxt, valid, mode):
        r = dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                 dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=mint.astype(F), tmax=maxt.astype(F),
                 valid=None if valid is None else valid.astype(np.uint8), width=W, height=H)
        if kind == 2:
            return orc.trace_raybvh(v9, *handoff, r, p, mode)
        return orc.trace(v9, r, p, mode)

    def dot(a, b):
        return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    h = trace(o, d, np.zeros(n, F), np.full(n, np.inf, F), None, orc.CLOSEST_PRIMARY)
    hit = h["id"] >= (1 if kind == 0 else 0)   # the naive layer drops hits on triangle 0
    tid, t = np.maximum(h["id"], 0), h["t"]
    L, beta = np.zeros(n, F), np.ones(n, F)
    Lc = F(albedo) * F(0.31830988618379067154)
    lt = np.array(light, F)
    for depth in range(max_depth + 1):
        pt = (o + d * np.where(hit, t, F(0))[:, None]).astype(F)
        eps = (F(1e-3) * np.where(hit, t, F(0))).astype(F)
        v = (lt[None, :] - pt).astype(F)
        dist = np.sqrt(dot(v, v))
        sd = (v * (F(1) / np.where(hit, dist, F(1)))[:, None]).astype(F)
        occ = trace(pt, sd, eps, dist, hit, orc.OCCLUDED)["occ"]
        nn = nrm[tid]
        nn = np.where((dot(nn, d) > 0)[:, None], -nn, nn).astype(F)
        c = np.maximum(F(0), dot(nn, sd))
        with np.errstate(divide="ignore", invalid="ignore"):
            add = (((beta * Lc) * F(intensity)) * c) / (dist * dist)
        L = np.where(hit & (occ == 0), L + add, L).astype(F)
        if depth >= max_depth or not hit.any():
            break
        j = np.arange(n, dtype=np.uint64)
        inner = _mix32(np.full(n, (depth + 0 * 977) & 0xffffffff, np.uint32))
        hh = _mix32(np.uint32(seed) ^ _mix32(((j * 0x9e3779b9) & 0xffffffff).astype(np.uint32) ^ inner))
        u1 = (hh >> np.uint32(8)).astype(F) * F(1.0 / 16777216.0)
        u2 = (_mix32(hh) >> np.uint32(8)).astype(F) * F(1.0 / 16777216.0)
        r = np.sqrt(u1)
        phi = F(6.28318530717958647692) * u2
        lx = r * np.array([cosf(float(x)) for x in phi], F)
        ly = r * np.array([sinf(float(x)) for x in phi], F)
        lz = np.sqrt(np.maximum(F(0), F(1) - u1))
        sgn = np.where(nn[:, 2] >= 0, F(1), F(-1))
        a = F(-1) / (sgn + nn[:, 2])
        b = nn[:, 0] * nn[:, 1] * a
        t1 = np.stack([F(1) + sgn * nn[:, 0] * nn[:, 0] * a, sgn * b, -sgn * nn[:, 0]], 1)
        t2 = np.stack([b, sgn + nn[:, 1] * nn[:, 1] * a, -nn[:, 1]], 1)
        wo = (t1 * lx[:, None] + t2 * ly[:, None]) + nn * lz[:, None]
        o = np.where(hit[:, None], pt, o).astype(F)
        d = np.where(hit[:, None], wo, d).astype(F)
        beta = np.where(hit, beta * F(albedo), beta).astype(F)
        h = trace(o, d, eps, np.full(n, np.inf, F), hit, orc.CLOSEST_SECONDARY)
        hit = hit & (h["id"] >= (1 if kind == 0 else 0))
        tid, t = np.maximum(h["id"], 0), h["t"]
    img = (L / F(spp)).reshape(yres, xres, spp)
    out = np.zeros((yres, xres), F)
    for k in range(spp):
        out += img[:, :, k]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("accel", ["naive", "rayhierarchy", "ray-bvh"])
def test_breadth_first_path_matches_oracle(scenes, orc, accel):
    """Path tracing with two bounces issues Intersect(..., bounce) and IntersectP batches; the
    rendered image equals the same loop run over the oracle's traces (naive: its own
    predicates and the layer's drop of triangle 0; rayhierarchy / ray-bvh with box_dir)."""
    v9 = scenes.random_triangles(600, seed=3)
    params = {} if accel == "naive" else {"node": "box_dir"}
    img, st = _render(accel, v9, 96, 64, spp=2, max_depth=2,
                      params=";".join(f"{k}={v}" for k, v in params.items()))
    assert st[5] > 0, "no bounce rays were traced"
    ref = _oracle_path_image(orc, v9, accel, params, 96, 64, 2, 2)
    assert np.count_nonzero(ref) > 500
    bad = int(np.count_nonzero(img != ref))
    print(f"[path {accel}] pixels {img.size}, lit {np.count_nonzero(ref)}, differing {bad}",
          flush=True)
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7)
