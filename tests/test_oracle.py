"""CPU tests that pin the oracle: closed forms, special cases, invariants, brute force.

The reference ships no tests or golden vectors for this path (its OpenCL kernels cannot run
here), so every oracle function is pinned by what the mathematics fixes, chosen so that a
plausible mistake (strict vs inclusive bound, dropped term, transposed operand, wrong tie
rule) fails a check.
"""
import math

import numpy as np
import pytest


def tri(*pts):
    return [float(x) for p in pts for x in p]


# ---------------------------------------------------------------- predicates (closed forms)
# Triangle in the plane z = 2 with dyadic coordinates: every FP32 operation below is exact.
T0 = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))


@pytest.mark.parametrize("kind", ["strict", "naive", "naive_p", "pbrt", "repair"])
def test_closed_form_hit(orc, kind):
    # ray from (1, 1, 0) along +z hits at t = 2 with b1 = 1/4 (along e1), b2 = 1/4 (along e2)
    ok, (t, b1, b2) = orc.mt(kind, (1, 1, 0), (0, 0, 1), T0)
    assert ok
    assert (t, b1, b2) == (2.0, 0.25, 0.25)


def test_closed_form_scaled(orc):
    # origin (2, 1, -6), direction (0, 0, 0.5): t = 16, b1 = 2/4, b2 = 1/4
    ok, (t, b1, b2) = orc.mt("strict", (2, 1, -6), (0, 0, 0.5), T0)
    assert ok and (t, b1, b2) == (16.0, 0.5, 0.25)


def test_edges_strict_vs_inclusive(orc):
    # b1 == 0 (on edge v1-v3): strict leaf test rejects, pbrt accepts
    assert not orc.mt("strict", (0, 1, 0), (0, 0, 1), T0)[0]
    assert orc.mt("pbrt", (0, 1, 0), (0, 0, 1), T0)[0]
    # b2 == 0 (on edge v1-v2)
    assert not orc.mt("strict", (1, 0, 0), (0, 0, 1), T0)[0]
    assert orc.mt("pbrt", (1, 0, 0), (0, 0, 1), T0)[0]
    # b1 + b2 == 1 (hypotenuse)
    assert not orc.mt("strict", (2, 2, 0), (0, 0, 1), T0)[0]
    assert orc.mt("pbrt", (2, 2, 0), (0, 0, 1), T0)[0]
    # vertex v2: b1 == 1
    assert not orc.mt("strict", (4, 0, 0), (0, 0, 1), T0)[0]


def test_parallel_ray_rejected(orc):
    # direction in the triangle's plane: det == 0 -> every variant rejects
    for kind in ("strict", "naive", "naive_p", "pbrt", "repair"):
        assert not orc.mt(kind, (1, 1, 2), (1, 0, 0), T0)[0], kind


def test_naive_det_threshold(orc):
    # det = e1 . (d x e2) = 16 * dz for this triangle: |det| < 0.002 is rejected by the naive
    # closest test (naive.cl:32) but not by its shadow test (naive.cl:213)
    d_small = (0.0, 0.0, 1e-4)  # det = 0.0016
    assert not orc.mt("naive", (1, 1, 0), d_small, T0)[0]
    assert orc.mt("naive_p", (1, 1, 0), d_small, T0)[0]
    assert orc.mt("strict", (1, 1, 0), d_small, T0)[0]
    assert orc.mt("naive", (1, 1, 0), (0.0, 0.0, 2e-4), T0)[0]  # det = 0.0032


def test_naive_barycentric_slack(orc):
    # b1 slightly negative (-0.002): naive closest accepts (bound -0.003), strict rejects,
    # naive shadow rejects (bound -0.001)
    o = (-0.008, 1, 0)  # b1 = -0.008/4 = -0.002
    assert orc.mt("naive", o, (0, 0, 1), T0)[0]
    assert not orc.mt("naive_p", o, (0, 0, 1), T0)[0]
    assert not orc.mt("strict", o, (0, 0, 1), T0)[0]
    o2 = (-0.016, 1, 0)  # b1 = -0.004 < -0.003
    assert not orc.mt("naive", o2, (0, 0, 1), T0)[0]


def test_repair_region(orc):
    # shrunk region: b1 >= 0.002, b2 >= 0.002, b1 + b2 <= 0.998
    assert not orc.mt("repair", (0.004, 1, 0), (0, 0, 1), T0)[0]   # b1 = 0.001
    assert orc.mt("repair", (0.016, 1, 0), (0, 0, 1), T0)[0]       # b1 = 0.004
    assert not orc.mt("repair", (2.0, 1.996, 0), (0, 0, 1), T0)[0]  # b1+b2 = 0.999


def test_pbrt_cross_is_double(orc):
    # pbrt's Cross rounds once from double (geometry.h:481-488); the kernels round each
    # product in FP32.  Random inputs: both must agree on hits up to rounding, and the pbrt
    # variant must differ in the last bits for some inputs (otherwise it would be untested).
    g = np.random.default_rng(7)
    diffs = 0
    for _ in range(2000):
        v = g.normal(size=9).astype(np.float32)
        o = g.normal(size=3).astype(np.float32) * 3
        d = g.normal(size=3).astype(np.float32)
        a = orc.mt("strict", o, d, v)
        b = orc.mt("pbrt", o, d, v)
        if a[0] and b[0] and a[1] != b[1]:
            diffs += 1
            assert abs(a[1][0] - b[1][0]) <= 1e-4 * max(1.0, abs(b[1][0]))
    assert diffs > 0


# ---------------------------------------------------------------- culls
def node16(dmin, dmax, omin, omax, valid=1.0):
    return [*dmin, valid, *dmax, 0, *omin, 0, *omax, 0]


def test_cull_tri_nonconservative_example(orc):
    # Normalising a convex combination can leave the box of the normalised vertex directions.
    # The triangle lies in the plane x = 0.6 with vertex directions whose x components are at
    # most 0.6, while the direction to its interior point (0.6, 0.45, 0.45) has x = 0.686.  No
    # component straddles 0, so intersectsTri (intersectsNode6DB.cl:66-99) rejects a node whose
    # rays point there although those rays hit the triangle.
    v = tri((0.6, 0.8, 0.0), (0.6, 0.0, 0.8), (0.6, 0.8, 0.8))
    d = np.array([0.6, 0.45, 0.45], np.float32)
    d = d / np.linalg.norm(d)
    n = node16(d - 0.001, d + 0.001, (0, 0, 0), (0, 0, 0))
    assert not orc.cull("tri", n, v9=v)
    ok, (t, _, _) = orc.mt("pbrt", (0, 0, 0), d, v)
    assert ok and abs(t * d[0] - 0.6) < 1e-6
    # the straddling-axis OR accepts the classic arc case: directions (0.6, +-0.8, 0)
    v2 = tri((0.6, 0.8, 0.0), (0.6, -0.8, 0.0), (0.6, 0.0, 0.01))
    n2 = node16((0.99, -0.01, -0.01), (1.0, 0.01, 0.01), (0, 0, 0), (0, 0, 0))
    assert orc.cull("tri", n2, v9=v2)


def test_cull_tri_accepts_contained(orc):
    v = tri((-1, -1, 5), (1, -1, 5), (0, 1, 5))
    n = node16((-0.01, -0.01, 0.99), (0.01, 0.01, 1.0), (0, 0, 0), (0, 0, 0))
    assert orc.cull("tri", n, v9=v)
    # a node looking away (-z) but overlapping the straddling x range is still accepted
    # (the OR of intersectsNode6DB.cl:91-96 ignores z) ...
    n2 = node16((-0.01, -0.01, -1.0), (0.01, 0.01, -0.99), (0, 0, 0), (0, 0, 0))
    assert orc.cull("tri", n2, v9=v)
    # ... while one whose x and y ranges miss the triangle's is rejected
    n3 = node16((0.5, 0.5, -0.71), (0.51, 0.51, -0.7), (0, 0, 0), (0, 0, 0))
    assert not orc.cull("tri", n3, v9=v)


def test_cull_straddle_or_logic(orc):
    # The straddling-axis OR of intersectsNode6DB.cl:56-63: a triangle whose direction box
    # straddles x = 0 is accepted by a node overlapping it in x even with disjoint y.
    v = tri((-0.5, -3, 5), (0.5, -3, 5), (0, -2.5, 5))
    n_y_far = node16((-0.01, 0.5, 0.8), (0.01, 0.6, 0.9), (0, 0, 0), (0, 0, 0))
    assert orc.cull("tri", n_y_far, v9=v)
    # same node shifted so x no longer overlaps: rejected
    n_x_far = node16((0.5, 0.5, 0.8), (0.6, 0.6, 0.9), (0, 0, 0), (0, 0, 0))
    assert not orc.cull("tri", n_x_far, v9=v)


def _ray_box_hit(o, d, bmin, bmax):
    tmin, tmax = -np.inf, np.inf
    for a in range(3):
        if d[a] == 0:
            if o[a] < bmin[a] or o[a] > bmax[a]:
                return False
            continue
        t0, t1 = (bmin[a] - o[a]) / d[a], (bmax[a] - o[a]) / d[a]
        tmin, tmax = max(tmin, min(t0, t1)), min(tmax, max(t0, t1))
    return tmin <= tmax


def test_cull_ia_conservative_random(orc):
    # IA is conservative: when it rejects, no ray line of the node meets the box.
    g = np.random.default_rng(11)
    rejected = 0
    for _ in range(400):
        o = g.uniform(-1, 1, size=(8, 3))
        d = g.normal(size=(8, 3)) * 0.05 + g.normal(size=3)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        o, d = o.astype(np.float32), d.astype(np.float32)
        n = node16(d.min(0), d.max(0), o.min(0), o.max(0))
        c = g.uniform(-6, 6, size=3)
        bmin, bmax = (c - 0.3).astype(np.float32), (c + 0.3).astype(np.float32)
        if not orc.cull("ia", n, bmin=bmin, bmax=bmax):
            rejected += 1
            for k in range(8):
                assert not _ray_box_hit(o[k].astype(np.float64), d[k].astype(np.float64), bmin, bmax)
    assert rejected > 50


def test_cull_node6_grows_by_origin_extent(orc):
    # intersectsNode Minkowski-grows the box by half the origin extent: a box just beside the
    # direction cone is accepted once the node's origin box is wide enough.
    bmin, bmax = [2.0, 0.5, 5.0], [2.2, 0.7, 5.2]
    narrow = node16((-0.01, -0.01, 0.99), (0.01, 0.01, 1.0), (0, 0, 0), (0, 0, 0))
    wide = node16((-0.01, -0.01, 0.99), (0.01, 0.01, 1.0), (-2.5, 0, 0), (2.5, 0, 0))
    assert not orc.cull("node6", narrow, bmin=bmin, bmax=bmax)
    assert orc.cull("node6", wide, bmin=bmin, bmax=bmax)


# ---------------------------------------------------------------- ray hierarchy build
def test_rh_levels_rule(orc):
    assert orc.rh_levels(512, 512) == [(128, 128), (64, 64), (32, 32)]
    # stops after a level with an odd dimension (rayhierarchy.cpp:438-442)
    assert orc.rh_levels(24, 24) == [(6, 6), (3, 3)]
    assert orc.rh_levels(1024, 256) == [(256, 64), (128, 32), (64, 16)]
    assert orc.rh_levels(16, 16, height=1) == [(4, 4)]
    # ceil grids for partial tiles
    assert orc.rh_levels(10, 6) == [(3, 2), (2, 1)]
    # the level count halves with floor as the reference does (11 -> 5 is odd: two levels),
    # the grids with ceil (6 parents also cover the 11th leaf column)
    assert orc.rh_levels(33, 17, cx=3, cy=5) == [(11, 4), (6, 2)]
    assert orc.rh_levels(52, 32) == [(13, 8), (7, 4), (4, 2)]  # 13 -> 6 -> 3
    # one leaf: the reference's floor gives 0 parents (nothing traced); stop above the leaves
    assert orc.rh_levels(16, 16, cx=16, cy=16) == [(1, 1), (1, 1)]
    assert orc.rh_well_formed(512, 512) and orc.rh_well_formed(1024, 256)
    assert not orc.rh_well_formed(10, 6)
    assert not orc.rh_well_formed(12, 16)  # 3 leaves wide: reference breaks above the leaves


def _brute_box(rays, x0, x1, y0, y1, use_valid):
    W = rays["width"]
    idx = [y * W + x for y in range(y0, y1) for x in range(x0, x1)]
    if use_valid and rays.get("valid") is not None:
        idx = [i for i in idx if rays["valid"][i]]
    if not idx:
        return None
    d = np.stack([rays["dx"][idx], rays["dy"][idx], rays["dz"][idx]], 1)
    o = np.stack([rays["ox"][idx], rays["oy"][idx], rays["oz"][idx]], 1)
    return d.min(0), d.max(0), o.min(0), o.max(0)


@pytest.mark.parametrize("W,H,use_valid", [(32, 32, False), (40, 24, True), (37, 21, True)])
def test_rh_build_brute_force(orc, scenes, W, H, use_valid):
    rays = scenes.random_rays(W * H, seed=W + H, width=W)
    g = np.random.default_rng(3)
    rays["valid"] = (g.random(W * H) < 0.7).astype(np.uint8)
    levels, nodes = orc.rh_build(rays, use_valid=use_valid)
    off = 0
    for l, (gw, gh) in enumerate(levels):
        tx, ty = 4 << l, 4 << l
        for j in range(gh):
            for k in range(gw):
                nd = nodes[off + j * gw + k]
                ref = _brute_box(rays, k * tx, min(W, (k + 1) * tx), j * ty, min(H, (j + 1) * ty),
                                 use_valid)
                if ref is None:
                    assert nd[3] == 0
                    continue
                assert nd[3] == 1
                np.testing.assert_array_equal(nd[0:3], ref[0])
                np.testing.assert_array_equal(nd[4:7], ref[1])
                np.testing.assert_array_equal(nd[8:11], ref[2])
                np.testing.assert_array_equal(nd[12:15], ref[3])
        off += gw * gh


# ---------------------------------------------------------------- traversal semantics
def _strict_brute(v9, rays, valid_only=False):
    """Nearest strict-predicate hit over all triangles (ties -> smallest id)."""
    n = rays["width"] * rays["height"]
    ntri = len(v9)
    ids = np.full(n, -1, np.int32)
    ts = np.full(n, np.inf, np.float32)
    from conftest import load_oracle
    orc = load_oracle()
    for i in range(n):
        if valid_only and rays.get("valid") is not None and not rays["valid"][i]:
            continue
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        for k in range(ntri):
            ok, (t, b1, b2) = orc.mt("strict", o, d, v9[k])
            if ok and t > rays["tmin"][i] and t < ts[i]:
                ts[i], ids[i] = t, k
    return ids, ts


def test_rh_ia_equals_brute_force(orc, scenes):
    # The IA cull is conservative, so with node "ia" the hierarchy must find exactly the
    # strict nearest hit of an exhaustive search (and the repair pass changes nothing here).
    v9 = scenes.random_triangles(48, seed=5, extent=1.0, size=0.3)
    rays = scenes.pinhole_rays(16, 16)
    got = orc.trace(v9, rays, orc.params("rayhierarchy", node="ia"), orc.CLOSEST_PRIMARY)
    ids, ts = _strict_brute(v9, rays)
    np.testing.assert_array_equal(got["id"], ids)
    np.testing.assert_array_equal(got["t"], ts)
    assert (ids >= 0).sum() > 20


def test_rh_boxdir_subset_of_brute_force(orc, scenes):
    # box_dir culls may drop candidates but never invent one: every reported hit is a strict
    # hit of that triangle, no closer than the exhaustive nearest.
    v9 = scenes.random_triangles(64, seed=6, extent=1.0, size=0.3)
    rays = scenes.pinhole_rays(16, 16)
    got = orc.trace(v9, rays, orc.params("rayhierarchy"), orc.CLOSEST_PRIMARY)
    ids, ts = _strict_brute(v9, rays)
    hit = got["id"] >= 0
    assert np.all(got["t"][hit] >= ts[hit])
    for i in np.nonzero(hit)[0]:
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        ok, (t, _, _) = orc.mt("strict", o, d, v9[got["id"][i]])
        assert ok and t == got["t"][i]


def test_rh_tie_smallest_id(orc):
    # two identical triangles: the nearest-hit reduction keeps the smaller id
    v = np.array([T0, T0], dtype=np.float32)
    rays = dict(ox=np.array([1, 1], np.float32), oy=np.array([1, 1.5], np.float32),
                oz=np.zeros(2, np.float32), dx=np.zeros(2, np.float32),
                dy=np.zeros(2, np.float32), dz=np.ones(2, np.float32),
                tmin=np.zeros(2, np.float32), tmax=np.full(2, np.inf, np.float32), valid=None,
                width=2, height=1)
    for kind in ("naive", "rayhierarchy"):
        got = orc.trace(v, rays, orc.params(kind, height=1), orc.CLOSEST_PRIMARY)
        assert list(got["id"]) == [0, 0], kind


def test_repair_pass_hit_at_tmin(orc):
    # A triangle through the ray origin gives t == tmin == 0: the main leaf test (t > tmin)
    # rejects it, the repair pass (t >= tmin, t <= tHit - EPS) takes it over a farther hit.
    plane = tri((-1, -1, 0), (2, -1, 0), (-1, 2, 0))
    far = tri((-1, -1, 5), (2, -1, 5), (-1, 2, 5))
    v = np.array([far, plane], dtype=np.float32)
    rays = dict(ox=np.zeros(1, np.float32), oy=np.zeros(1, np.float32), oz=np.zeros(1, np.float32),
                dx=np.zeros(1, np.float32), dy=np.zeros(1, np.float32), dz=np.ones(1, np.float32),
                tmin=np.zeros(1, np.float32), tmax=np.full(1, np.inf, np.float32), valid=None,
                width=1, height=1)
    p = orc.params("rayhierarchy", height=1)
    got = orc.trace(v, rays, p, orc.CLOSEST_PRIMARY)
    assert got["id"][0] == 1 and got["t"][0] == 0.0
    assert got["stats"].repair_rays == 1
    p0 = orc.params("rayhierarchy", height=1, repair_run=0)
    got0 = orc.trace(v, rays, p0, orc.CLOSEST_PRIMARY)
    assert got0["id"][0] == 0 and got0["t"][0] == 5.0


def test_secondary_validity(orc, scenes):
    v9 = scenes.random_triangles(64, seed=8, extent=1.0, size=0.3)
    rays = scenes.pinhole_rays(16, 16)
    rays["valid"] = (np.arange(256) % 3 != 0).astype(np.uint8)
    got = orc.trace(v9, rays, orc.params("rayhierarchy", node="ia"), orc.CLOSEST_SECONDARY)
    assert np.all(got["id"][rays["valid"] == 0] == -1)
    ids, ts = _strict_brute(v9, rays, valid_only=True)
    np.testing.assert_array_equal(got["id"], ids)


def test_occlusion_ia_equals_brute_force(orc, scenes):
    v9 = scenes.random_triangles(64, seed=9, extent=1.0, size=0.3)
    rays = scenes.random_rays(256, seed=4, lo=-1.5, hi=1.5, maxt=2.0, width=16)
    got = orc.trace(v9, rays, orc.params("rayhierarchy", node="ia"), orc.OCCLUDED)
    exp = np.zeros(256, np.uint8)
    for i in range(256):
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        for k in range(len(v9)):
            ok, (t, _, _) = orc.mt("strict", o, d, v9[k])
            if ok and rays["tmin"][i] < t < rays["tmax"][i]:
                exp[i] = 1
                break
    np.testing.assert_array_equal(got["occ"], exp)
    assert 10 < exp.sum() < 246


def test_naive_matches_bruteforce_rule(orc, scenes):
    v9 = scenes.random_triangles(32, seed=10, extent=1.0, size=0.3)
    rays = scenes.pinhole_rays(12, 12)
    got = orc.trace(v9, rays, orc.params("naive"), orc.CLOSEST_PRIMARY)
    for i in range(rays["width"] * rays["height"]):
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        best = (np.inf, -1)
        for k in range(len(v9)):
            ok, (t, _, _) = orc.mt("naive", o, d, v9[k])
            if ok and t >= rays["tmin"][i] and (best[1] < 0 or t < best[0]):
                best = (t, k)
        assert got["id"][i] == best[1]


# ---------------------------------------------------------------- shading record
def test_shading_closed_form(orc, scenes):
    # default UVs (0,0),(1,0),(1,1) give u = b1 + b2, v = b2 (computeDpTuTv.cl:87-89);
    # dpdu = p2 - p1 and dpdv = p3 - p2 follow from the UV determinant 1
    v9 = np.array([T0], np.float32)
    rays = dict(ox=np.array([1.0], np.float32), oy=np.array([2.0], np.float32),
                oz=np.zeros(1, np.float32), dx=np.zeros(1, np.float32), dy=np.zeros(1, np.float32),
                dz=np.ones(1, np.float32), tmin=np.zeros(1, np.float32),
                tmax=np.full(1, np.inf, np.float32), valid=None, width=1, height=1)
    s = orc.shading(v9, None, rays, np.array([0], np.int32))
    assert s["u"][0] == 0.75 and s["v"][0] == 0.5
    np.testing.assert_array_equal(s["dpdu"], [4, 0, 0])
    np.testing.assert_array_equal(s["dpdv"], [-4, 4, 0])
    # degenerate UVs: the coordinate-system fallback gives an orthonormal pair around n
    uv = np.zeros(6, np.float32)
    s2 = orc.shading(v9, uv, rays, np.array([0], np.int32))
    du, dv = s2["dpdu"], s2["dpdv"]
    assert abs(np.dot(du, dv)) < 1e-6 and abs(np.linalg.norm(du) - 1) < 1e-6
    assert abs(abs(dv[0] * 0 + dv[1] * 0) + 0) < 1e-6 and abs(np.linalg.norm(dv) - 1) < 1e-6
    # UVs (0,0), (1,0), (1,d) have determinant d exactly (du1 = -1, du2 = 0, dv1 = dv2 = -d).
    # The naive kernels fall back for |det| < EPS = 0.002 (naive.cl:60-61), computeDpTuTv only
    # for det == 0: with d = +-0.001 the two rules differ, at d = 0.002 (not below EPS) and
    # d = 0.004 they agree.
    for d, differ in ((0.001, True), (-0.001, True), (0.002, False), (0.004, False)):
        uvd = np.array([0, 0, 1, 0, 1, d], np.float32)
        a = orc.shading(v9, uvd, rays, np.array([0], np.int32))
        b = orc.shading(v9, uvd, rays, np.array([0], np.int32), naive=True)
        assert (not np.array_equal(a["dpdv"], b["dpdv"])) == differ, d
        if differ:
            np.testing.assert_array_equal(b["dpdu"], s2["dpdu"])   # the fallback pair
            np.testing.assert_array_equal(b["dpdv"], s2["dpdv"])
            assert np.linalg.norm(a["dpdv"]) > 100                 # -dp2 / det
        np.testing.assert_array_equal(a["u"], b["u"])              # u, v do not depend on it


# ---------------------------------------------------------------- BVH build / handoff
def _check_handoff(nodes, roots, order, ntri, bvh_height):
    assert sorted(order.tolist()) == list(range(ntri))
    covered = np.zeros(ntri, np.int32)
    for r in roots:
        stack = [(int(r), 0)]
        while stack:
            n, depth = stack.pop()
            assert depth <= bvh_height
            nd = nodes[n]
            if nd["n_primitives"]:
                covered[nd["prim_offset"]:nd["prim_offset"] + nd["n_primitives"]] += 1
            else:
                a = int(nd["prim_offset"])
                for c in (a, a + 1):
                    ch = nodes[c]
                    for lo, hi in (("ax", "bx"), ("ay", "by"), ("az", "bz")):
                        assert nd[lo] <= ch[lo] and ch[hi] <= nd[hi]
                    stack.append((c, depth + 1))
    assert np.all(covered == 1), "frontier must cover every primitive exactly once"


def test_bvh_handoff_invariants(orc, scenes):
    v9 = scenes.random_triangles(3000, seed=12)
    nodes, roots, order = orc.raybvh_handoff(v9)
    _check_handoff(nodes, roots, order, 3000, 3)
    # leaf boxes contain their triangles
    for nd in nodes[nodes["n_primitives"] > 0][:200]:
        t = v9[order[nd["prim_offset"]]].reshape(3, 3)
        assert np.all(t.min(0) >= [nd["ax"], nd["ay"], nd["az"]])
        assert np.all(t.max(0) <= [nd["bx"], nd["by"], nd["bz"]])


def test_bvh_handoff_parts(orc, scenes):
    # more than 60,000 triangles -> equal parts of ceil(n/parts), each with its own BVH
    v9 = scenes.cluster_scene(61000, seed=13)
    nodes, roots, order = orc.raybvh_handoff(v9)
    _check_handoff(nodes, roots, order, 61000, 3)
    leaf = nodes[nodes["n_primitives"] > 0]
    assert leaf["prim_offset"].min() == 0 and leaf["prim_offset"].max() == 60999


def test_bvh_sah_leaf_sizes(orc, scenes):
    v9 = scenes.random_triangles(500, seed=14)
    nodes, top, order = orc.bvh_build(v9, max_prims=4, gpu=True, bvh_height=64)
    assert top == 1  # nothing skipped: the root is the only frontier node
    # maxnodeprims 4 with the SAH split of this fork: every leaf holds one primitive unless
    # centroids coincide
    assert set(nodes["n_primitives"][nodes["n_primitives"] > 0].tolist()) == {1}
    assert nodes.size == 2 * 500 - 1


def test_cpu_bvh_matches_exhaustive(orc, scenes):
    v9 = scenes.random_triangles(800, seed=15)
    rays = scenes.pinhole_rays(48, 48)
    bvh = orc.CpuBvh(v9)
    got = bvh.closest(rays, nthreads=2)
    ex = orc.exact_closest(v9, rays)
    np.testing.assert_array_equal(got["t"], ex["t"])
    assert (got["id"] == ex["id"]).mean() > 0.999
    bvh.close()


# ---------------------------------------------------------------- ray-bvh traversal pins
def _strict_occ_brute(v9, rays):
    n = rays["width"] * rays["height"]
    exp = np.zeros(n, np.uint8)
    for i in range(n):
        if rays.get("valid") is not None and not rays["valid"][i]:
            continue
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        for k in range(len(v9)):
            ok, (t, _, _) = orc_mt_strict(o, d, v9[k])
            if ok and rays["tmin"][i] < t < rays["tmax"][i]:
                exp[i] = 1
                break
    return exp


def orc_mt_strict(o, d, v):
    from conftest import load_oracle
    return load_oracle().mt("strict", o, d, v)


def test_raybvh_ia_equals_brute_force(orc, scenes):
    # With node "ia" every cull of the lockstep chain is the conservative interval test
    # against a BVH box that contains its triangles, so ray-bvh must return the exhaustive
    # strict nearest hit (primary, and secondary with one primitive per leaf) and the
    # exhaustive occlusion.
    v9 = scenes.random_triangles(60, seed=41, extent=1.0, size=0.3)
    nodes, roots, order = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3)
    p = orc.params("ray-bvh", node="ia")
    rays = scenes.pinhole_rays(16, 16)
    got = orc.trace_raybvh(v9, nodes, roots, order, rays, p, orc.CLOSEST_PRIMARY)
    ids, ts = _strict_brute(v9, rays)
    np.testing.assert_array_equal(got["id"], ids)
    np.testing.assert_array_equal(got["t"], ts)
    assert (ids >= 0).sum() > 20
    sec = scenes.random_rays(256, seed=42, lo=-1.5, hi=1.5, maxt=np.inf, width=16)
    sec["valid"] = (np.arange(256) % 5 != 0).astype(np.uint8)
    got = orc.trace_raybvh(v9, nodes, roots, order, sec, p, orc.CLOSEST_SECONDARY)
    ids, ts = _strict_brute(v9, sec, valid_only=True)
    np.testing.assert_array_equal(got["id"], ids)
    assert (ids >= 0).sum() > 10
    sh = scenes.random_rays(256, seed=43, lo=-1.5, hi=1.5, maxt=2.0, width=16)
    got = orc.trace_raybvh(v9, nodes, roots, order, sh, p, orc.OCCLUDED)
    exp = _strict_occ_brute(v9, sh)
    np.testing.assert_array_equal(got["occ"], exp)
    assert 10 < exp.sum() < 246


def _paired_scene(n, seed):
    """n triangles plus, for each, its point reflection through its bounding-box centre: both
    have the same BVH centroid bit for bit (dyadic coordinates), so SAH leaves hold pairs."""
    g = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = g.integers(-256, 256, size=3) / 128.0
        off = g.integers(-24, 25, size=(3, 3)) / 256.0
        off -= (off.min(0) + off.max(0)) / 2  # bounding-box centre at c exactly
        a = c + off
        b = c - off
        out.append(a.reshape(-1))
        out.append(b.reshape(-1))
    return np.array(out, np.float32)


def test_raybvh_secondary_first_primitive_rule(orc, scenes):
    # intersection6DB_BVH2.cl tests only the first primitive of each BVH leaf for secondary
    # rays.  With maxBVHPrim 4 and coincident centroids the leaves hold two triangles: the
    # secondary result is the exhaustive strict nearest hit over first primitives only (IA
    # keeps every candidate), while primary rays see every primitive.
    v9 = _paired_scene(40, seed=44)
    nodes, roots, order = orc.raybvh_handoff(v9, max_prims=4, bvh_height=3)
    leaves = nodes[nodes["n_primitives"] > 0]
    assert (leaves["n_primitives"] >= 2).sum() >= 20
    first = np.array(sorted(int(order[lf["prim_offset"]]) for lf in leaves))
    p = orc.params("ray-bvh", node="ia")
    rays = scenes.random_rays(256, seed=45, lo=-2.0, hi=2.0, maxt=np.inf, width=16)
    got = orc.trace_raybvh(v9, nodes, roots, order, rays, p, orc.CLOSEST_SECONDARY)
    sub_ids, sub_t = _strict_brute(v9[first], rays, valid_only=True)
    exp = np.where(sub_ids >= 0, first[np.maximum(sub_ids, 0)], -1)
    np.testing.assert_array_equal(got["id"], exp)
    all_ids, _ = _strict_brute(v9, rays)
    assert (all_ids != exp).sum() > 0, "the rule must change some result here"
    got0 = orc.trace_raybvh(v9, nodes, roots, order, rays, p, orc.CLOSEST_PRIMARY)
    np.testing.assert_array_equal(got0["id"], all_ids)


def _rays_of(o, d, W, H, tmax=np.inf):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = W * H
    return dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, np.float32),
                tmax=np.full(n, tmax, np.float32), valid=None, width=W, height=H)


def test_rh_chain_root_cull_rejects_leaf_pass(orc):
    # Hand-built: 2x2 rays, 1x1 leaves, height 2 (four leaves under one root).  Ray 0 starts at
    # the origin and points along +z at triangle A (z = 5 around the axis): its leaf accepts A
    # (the vertex directions straddle x and y = 0).  The other three rays start at
    # (100, 100, 0), so the root's origin box is [0,100]^2 x {0} with centre (50, 50, 0);
    # seen from there A's vertex directions have x, y in (-0.72, -0.69) -- no straddling axis
    # and no overlap with the root's direction box (0, 0, 1) -- so intersectsTri rejects A at
    # the root and ray 0 misses although it hits A (intersection6DB.cl:51-119 culls at every
    # level from the top).
    A = tri((-1, -1, 5), (1, -1, 5), (0, 1, 5))
    v = np.array([A], np.float32)
    o = [(0, 0, 0), (100, 100, 0), (100, 100, 0), (100, 100, 0)]
    d = [(0, 0, 1)] * 4
    rays = _rays_of(o, d, 2, 2)
    assert orc.mt("strict", (0, 0, 0), (0, 0, 1), A)[0]
    leaf = node16((0, 0, 1), (0, 0, 1), (0, 0, 0), (0, 0, 0))
    root = node16((0, 0, 1), (0, 0, 1), (0, 0, 0), (100, 100, 0))
    assert orc.cull("tri", leaf, v9=A)
    assert not orc.cull("tri", root, v9=A)
    p = orc.params("rayhierarchy", cx=1, cy=1, height=2, repair_run=0)
    got = orc.trace(v, rays, p, orc.CLOSEST_PRIMARY)
    assert got["id"][0] == -1
    # the same ray alone (its own root) finds A
    alone = _rays_of([(0, 0, 0)], [(0, 0, 1)], 1, 1)
    assert orc.trace(v, alone, p, orc.CLOSEST_PRIMARY)["id"][0] == 0


def test_rh_chain_leaf_cull_rejects_root_pass(orc):
    # Hand-built: the non-conservative example of intersectsTri at the leaf.  Ray 0 points at
    # the interior point (0.6, 0.45, 0.45) of a triangle in the plane x = 0.6; its direction's
    # x component 0.687 lies above every vertex direction's (at most 0.6), so its 1x1 leaf
    # rejects the triangle.  The other rays point at the vertices, so the root's direction box
    # covers the triangle's and the root accepts it: the ray misses because of the leaf level.
    B = tri((0.6, 0.8, 0.0), (0.6, 0.0, 0.8), (0.6, 0.8, 0.8))
    v = np.array([B], np.float32)
    dirs = np.array([(0.6, 0.45, 0.45), (0.6, 0.8, 0.0), (0.6, 0.0, 0.8), (0.6, 0.8, 0.8)],
                    np.float64)
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    rays = _rays_of([(0, 0, 0)] * 4, dirs, 2, 2)
    assert orc.mt("strict", (0, 0, 0), dirs[0], B)[0]
    lo, hi = dirs.min(0), dirs.max(0)
    assert orc.cull("tri", node16(lo, hi, (0, 0, 0), (0, 0, 0)), v9=B)
    assert not orc.cull("tri", node16(dirs[0], dirs[0], (0, 0, 0), (0, 0, 0)), v9=B)
    p = orc.params("rayhierarchy", cx=1, cy=1, height=2, repair_run=0)
    got = orc.trace(v, rays, p, orc.CLOSEST_PRIMARY)
    assert got["id"][0] == -1
    # with 2x2 leaves (one leaf holding all four rays) the leaf is the root: the ray hits
    p1 = orc.params("rayhierarchy", cx=2, cy=2, height=1, repair_run=0)
    assert orc.trace(v, rays, p1, orc.CLOSEST_PRIMARY)["id"][0] == 0


def _node_at(levels, nodes, l, k, j):
    off = sum(w * h for w, h in levels[:l])
    return nodes[off + j * levels[l][0] + k]


def _flat_result(orc, v9, rays, mode, cx, cy, candidate):
    """Nearest strict hit / occlusion over an explicit candidate set candidate[(k, j, tri)]."""
    W, H = rays["width"], rays["height"]
    n = W * H
    ids = np.full(n, -1, np.int32)
    ts = np.full(n, np.inf, np.float32)
    occ = np.zeros(n, np.uint8)
    for i in range(n):
        if mode != orc.CLOSEST_PRIMARY and rays.get("valid") is not None and not rays["valid"][i]:
            continue
        x, y = i % W, i // W
        o = (rays["ox"][i], rays["oy"][i], rays["oz"][i])
        dd = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        for k in range(len(v9)):
            if not candidate.get((x // cx, y // cy, k), False):
                continue
            ok, (t, _, _) = orc.mt("strict", o, dd, v9[k])
            if not ok:
                continue
            if mode == orc.OCCLUDED:
                if rays["tmin"][i] < t < rays["tmax"][i]:
                    occ[i] = 1
            elif t > rays["tmin"][i] and t < ts[i]:
                ts[i], ids[i] = t, k
    return ids, ts, occ


def _wide_rays(W, H, seed):
    # origins spread over a box and directions over a cone, so that nodes of different levels
    # see a triangle from different origin-box centres
    g = np.random.default_rng(seed)
    n = W * H
    o = g.uniform(-1.0, 1.0, size=(n, 3)) * [1.0, 1.0, 0.2] + [0, 0, -3]
    d = g.normal(size=(n, 3)) * 0.35 + [0, 0, 1]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays_of(o.astype(np.float32), d.astype(np.float32), W, H)


def _with_valid(rays, seed):
    rays["valid"] = (np.random.default_rng(seed).random(rays["width"] * rays["height"])
                     < 0.8).astype(np.uint8)
    return rays


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_rh_boxdir_chain_flat_restatement(orc, mode):
    # rayhierarchy box_dir: a triangle is tested at a leaf iff every node from the top level
    # down to that leaf accepts it (intersectsTri for primary rays, intersectsNode otherwise;
    # intersection6DB.cl:51-119 and variants), and the traversal evaluates a node's cull iff
    # every ancestor accepted (and, for secondary and shadow rays, the node holds a valid
    # ray).  Both are evaluated here flat, per (triangle, node), with the single-node
    # predicates pinned above: the oracle's depth-first traversal must give the same hits and
    # the same cull and leaf-pair counts.  Primary rays with spread origins must also give
    # pairs that only an interior level rejects and pairs that only the leaf rejects.
    v9 = np.array([tri((-1, -1, 5), (1, -1, 5), (0, 1, 5))], np.float32)
    g = np.random.default_rng(46)
    v9 = np.concatenate([v9, (g.uniform(-1.2, 1.2, size=(39, 1, 3)) + [0, 0, 1.5]
                              + g.uniform(-0.4, 0.4, size=(39, 3, 3))).reshape(-1, 9)
                        .astype(np.float32)])
    W = H = 8
    rays = _wide_rays(W, H, seed=47)
    if mode != 0:
        rays = _with_valid(rays, 7)
    if mode == 2:
        rays["tmax"] = np.full(W * H, 4.0, np.float32)
    cx = cy = 2
    use_valid = mode != 0
    levels, nodes = orc.rh_build(rays, cx=cx, cy=cy, height=3, use_valid=use_valid)
    L = len(levels)

    def cull(nd, k):
        if mode == 0:
            return orc.cull("tri", nd, v9=v9[k])
        t = v9[k].reshape(3, 3)
        return orc.cull("node6", nd, bmin=t.min(0), bmax=t.max(0))

    culls = pairs = interior_only = leaf_only = 0
    candidate = {}
    for k in range(len(v9)):
        passed = {}
        for l in range(L - 1, -1, -1):
            gw, gh = levels[l]
            for j in range(gh):
                for kk in range(gw):
                    above = l == L - 1 or passed.get((l + 1, kk >> 1, j >> 1), False)
                    nd = _node_at(levels, nodes, l, kk, j)
                    if not above or (use_valid and nd[3] == 0):
                        continue
                    culls += 1
                    passed[(l, kk, j)] = cull(nd, k)
        for j in range(levels[0][1]):
            for kk in range(levels[0][0]):
                ok = passed.get((0, kk, j), False)
                candidate[(kk, j, k)] = ok
                pairs += ok
                nd0 = _node_at(levels, nodes, 0, kk, j)
                if use_valid and nd0[3] == 0:
                    continue
                chain = [cull(_node_at(levels, nodes, l, kk >> l, j >> l), k) for l in range(L)]
                interior_only += chain[0] and not all(chain[1:])
                leaf_only += (not chain[0]) and all(chain[1:])
    ids, ts, occ = _flat_result(orc, v9, rays, mode, cx, cy, candidate)
    p = orc.params("rayhierarchy", cx=cx, cy=cy, height=3, repair_run=0)
    got = orc.trace(v9, rays, p, mode)
    if mode == 2:
        np.testing.assert_array_equal(got["occ"], occ)
    else:
        np.testing.assert_array_equal(got["id"], ids)
        np.testing.assert_array_equal(got["t"], ts)
    assert got["stats"].cull_tests == culls
    assert got["stats"].leaf_pairs == pairs
    if mode == 0:
        assert interior_only > 0 and leaf_only > 0, (interior_only, leaf_only)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_raybvh_lockstep_chain_flat_restatement(orc, scenes, mode):
    # ray-bvh box_dir: the product traversal (intersection6DB_BVH.cl:51-223) visits the pairs
    # (hierarchy node R, BVH node B) step by step from (top node, frontier root): step s pairs
    # the hierarchy node of level max(top - s, 0) with the BVH node of depth min(s, d) (d the
    # depth of B's leaf below its root), so a pair's parent is the pair one step up, and a pair
    # is visited (and culled with intersectsNode) iff its parent was visited and accepted (and
    # R holds a valid ray for secondary and shadow rays).  Here the visited set is grown
    # breadth first from those rules and the (hierarchy leaf, BVH leaf) pairs reached set the
    # candidates; the oracle's recursive traversal must give the same hits and counts.
    v9 = scenes.random_triangles(48, seed=48, extent=1.0, size=0.4)
    nodes, roots, order = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3)
    W = H = 8
    rays = _wide_rays(W, H, seed=49)
    if mode != 0:
        rays = _with_valid(rays, 8)
    if mode == 2:
        rays["tmax"] = np.full(W * H, 4.0, np.float32)
    cx = cy = 2
    use_valid = mode != 0
    levels, rnodes = orc.rh_build(rays, cx=cx, cy=cy, height=3, use_valid=use_valid)
    top = len(levels) - 1

    def children(l, k, j, b):
        """Pairs one step below (l, k, j, b)."""
        nb = nodes[b]
        rk = [(l - 1, ck, cj) for cj in (2 * j, 2 * j + 1) for ck in (2 * k, 2 * k + 1)
              if l > 0 and ck < levels[l - 1][0] and cj < levels[l - 1][1]] if l > 0 else [(l, k, j)]
        bk = [b] if nb["n_primitives"] else [int(nb["prim_offset"]), int(nb["prim_offset"]) + 1]
        return [(r[0], r[1], r[2], c) for r in rk for c in bk]

    culls = 0
    candidate = {}
    frontier = [(top, k, j, int(r)) for r in roots for j in range(levels[top][1])
                for k in range(levels[top][0])]
    while frontier:
        nxt = []
        for (l, k, j, b) in frontier:
            nd = _node_at(levels, rnodes, l, k, j)
            if use_valid and nd[3] == 0:
                continue
            culls += 1
            box = nodes[b]
            if not orc.cull("node6", nd, bmin=(box["ax"], box["ay"], box["az"]),
                            bmax=(box["bx"], box["by"], box["bz"])):
                continue
            if l == 0 and box["n_primitives"]:
                cnt = 1 if mode == 1 else int(box["n_primitives"])
                for f in range(cnt):
                    candidate[(k, j, int(order[box["prim_offset"] + f]))] = True
                continue
            nxt.extend(children(l, k, j, b))
        frontier = nxt
    ids, ts, occ = _flat_result(orc, v9, rays, mode, cx, cy, candidate)
    p = orc.params("ray-bvh", cx=cx, cy=cy, height=3)
    got = orc.trace_raybvh(v9, nodes, roots, order, rays, p, mode)
    if mode == 2:
        np.testing.assert_array_equal(got["occ"], occ)
    else:
        np.testing.assert_array_equal(got["id"], ids)
        np.testing.assert_array_equal(got["t"], ts)
    assert got["stats"].cull_tests == culls
    assert (ids >= 0).sum() > 5 if mode != 2 else occ.sum() > 5


# ---------------------------------------------------------------- parity bookkeeping flags
def test_flags_near_tie(orc):
    # two triangles parallel to T0 at z = 2 and z = 2 * (1 + d): the ray along +z from z = 0
    # sees both; d = 2^-21 (5e-7) is a near tie (within 1e-6 * t), d = 2^-14 is not
    def scene(dz):
        far = tri((0, 0, 2 + 2 * dz), (4, 0, 2 + 2 * dz), (0, 4, 2 + 2 * dz))
        return np.array([T0, far], np.float32)
    rays = _rays_of([(1, 1, 0)], [(0, 0, 1)], 1, 1)
    p = orc.params("rayhierarchy", height=1)
    got = orc.trace(scene(2.0 ** -21), rays, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] & orc.NEAR_TIE
    got = orc.trace(scene(2.0 ** -14), rays, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and not got["flags"][0] & orc.NEAR_TIE
    # identical triangles: an exact tie
    got = orc.trace(np.array([T0, T0], np.float32), rays, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] & orc.NEAR_TIE


def test_flags_ulp_marginal(orc):
    # Direction (1/16, 1/16, 1) keeps every operand dyadic.  The ray from (-1/8, 1, 0) meets
    # T0 at (0, 9/8, 2), on the edge b1 == 0: the rejection rests on a numerator that is
    # exactly 0, so the ray is marginal (any other rounding could accept it).  The ray from
    # (1, 1, 0) meets the interior (b1 = b2 = 9/32): not marginal; neither is the edge ray
    # behind a robust nearer hit.
    d = (1 / 16, 1 / 16, 1)
    p = orc.params("rayhierarchy", node="ia", height=1)
    edge = _rays_of([(-1 / 8, 1, 0)], [d], 1, 1)
    assert not orc.mt("strict", (-1 / 8, 1, 0), d, T0)[0]
    got = orc.trace(np.array([T0], np.float32), edge, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == -1 and got["flags"][0] & orc.ULP_MARGINAL
    inner = _rays_of([(1, 1, 0)], [d], 1, 1)
    got = orc.trace(np.array([T0], np.float32), inner, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] == 0
    near = tri((-8, -8, 1), (8, -8, 1), (0, 8, 1))  # robust hit at t = 1 in front of T0
    got = orc.trace(np.array([near, T0], np.float32), edge, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] == 0
    # occlusion: the edge ray is marginal unless something else occludes it robustly
    sh = _rays_of([(-1 / 8, 1, 0)], [d], 1, 1, tmax=10.0)
    got = orc.trace(np.array([T0], np.float32), sh, p, orc.OCCLUDED, flags=True)
    assert got["occ"][0] == 0 and got["flags"][0] & orc.ULP_MARGINAL
    got = orc.trace(np.array([near, T0], np.float32), sh, p, orc.OCCLUDED, flags=True)
    assert got["occ"][0] == 1 and got["flags"][0] == 0
    # a hit exactly at tmax is a marginal occlusion decision
    at = _rays_of([(1, 1, 0)], [d], 1, 1, tmax=2.0)
    got = orc.trace(np.array([T0], np.float32), at, p, orc.OCCLUDED, flags=True)
    assert got["occ"][0] == 0 and got["flags"][0] & orc.ULP_MARGINAL


def test_flags_marginal_cull_descends(orc):
    # The IA cull of a node whose direction interval touches 0 exactly (d.x == 0) is decided
    # on its threshold: a hit below it is marginal even though the cull accepts.  With d.x =
    # 1/16 the straddle test is robust.
    v = np.array([T0], np.float32)
    p = orc.params("rayhierarchy", node="ia", height=1)
    r0 = _rays_of([(1, 1, 0)], [(1 / 16, 1 / 4, 1)], 1, 1)
    got = orc.trace(v, r0, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] == 0
    r1 = _rays_of([(1, 1, 0)], [(0, 1 / 4, 1)], 1, 1)
    got = orc.trace(v, r1, p, orc.CLOSEST_PRIMARY, flags=True)
    assert got["id"][0] == 0 and got["flags"][0] & orc.ULP_MARGINAL


def test_flags_do_not_change_results(orc, scenes):
    v9 = scenes.random_triangles(200, seed=50, extent=1.0, size=0.3)
    rays = scenes.pinhole_rays(24, 24)
    for accel in ("rayhierarchy", "ray-bvh"):
        p = orc.params(accel)
        if accel == "ray-bvh":
            nodes, roots, order = orc.raybvh_handoff(v9)
            run = lambda r, m, f: orc.trace_raybvh(v9, nodes, roots, order, r, p, m, flags=f)
        else:
            run = lambda r, m, f: orc.trace(v9, r, p, m, flags=f)
        a, b = run(rays, orc.CLOSEST_PRIMARY, False), run(rays, orc.CLOSEST_PRIMARY, True)
        np.testing.assert_array_equal(a["id"], b["id"])
        np.testing.assert_array_equal(a["t"], b["t"])
        assert a["stats"].cull_tests == b["stats"].cull_tests
        assert a["stats"].leaf_pairs == b["stats"].leaf_pairs


# ---------------------------------------------------------------- box_uv (5DB)
# Node "box_uv" stores the slope d.y/d.x and d.z intervals of its rays.  As C parses
# intersectsNode5DB.cl:79-81, a node passes when some corner direction of the origin-grown box
# has y >= 0, else when the corners' z range meets [dz min, dz max] within EPS = 0.002; the
# slope intervals never decide.  A point box seen from a point origin has one direction:
# (0, -3, 4) normalises to (0, -0.6, 0.8).
def uv16(slope, dz, omin=(0, 0, 0), omax=(0, 0, 0)):
    return node16((slope[0], dz[0], 0), (slope[1], dz[1], 0), omin, omax)


def test_cull_5db_z_interval_and_eps(orc):
    p = (0.0, -3.0, 4.0)
    cull = lambda dz, slope=(-1, 1): orc.cull("5db", uv16(slope, dz), bmin=p, bmax=p)
    assert cull((0.70, 0.90))
    assert cull((0.70, 0.7985))      # 0.8 < 0.7985 + EPS
    assert not cull((0.70, 0.797))   # 0.8 >= 0.797 + EPS
    assert cull((0.8005, 0.90))      # 0.8005 < 0.8 + EPS
    assert not cull((0.803, 0.90))   # EPS widens the upper ends only
    # the slope intervals never decide: slope 0 of this direction lies outside [100, 200]
    assert cull((0.70, 0.90), slope=(100, 200))
    assert not cull((-0.9, -0.7), slope=(-1e9, 1e9))


def test_cull_5db_y_sign_rule(orc):
    far = ((0, 0, 0), (5, 6))  # slope, dz: no corner direction has z near 5
    for p in ((0.0, 0.0, 5.0), (0.0, 1e-3, -5.0), (3.0, 2.0, 1.0)):
        assert orc.cull("5db", uv16(*far), bmin=p, bmax=p), p   # y >= 0: passes
    for p in ((0.0, -1e-3, 5.0), (3.0, -2.0, 1.0)):
        assert not orc.cull("5db", uv16(*far), bmin=p, bmax=p), p
    # one corner of a box with y >= 0 is enough
    assert orc.cull("5db", uv16(*far), bmin=(1, -2, 1), bmax=(2, 0, 2))
    assert not orc.cull("5db", uv16(*far), bmin=(1, -2, 1), bmax=(2, -1e-3, 2))


def test_cull_5db_grows_by_origin_extent(orc):
    # origin box [-1, 1]^3 (centre 0, half extent 1): the triangle box is grown by 1
    nd = lambda dz: uv16((0, 0), dz, omin=(-1, -1, -1), omax=(1, 1, 1))
    assert orc.cull("5db", nd((5, 6)), bmin=(0, -0.5, 4), bmax=(0, -0.5, 4))    # grown y reaches 0.5
    assert not orc.cull("5db", nd((5, 6)), bmin=(0, -2.5, 4), bmax=(0, -2.5, 4))
    assert orc.cull("5db", nd((-1, 1)), bmin=(0, -2.5, 4), bmax=(0, -2.5, 4))


def _row_rays(d, W, H, valid=None):
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = W * H
    r = _rays_of(np.zeros((n, 3)), d, W, H)
    if valid is not None:
        r["valid"] = np.asarray(valid, np.uint8)
    return r


def test_rh_build_5db_leaf(orc):
    # one 4x4 tile holding a 4x1 image: slopes 2, (d.x = 0), 0.5, -1; d.z 0.5, 0.3, 0.1, 0.7
    d = [(1, 2, 0.5), (0, 1, 0.3), (4, 2, 0.1), (2, -2, 0.7)]
    rays = _row_rays(d, 4, 1)
    _, nodes = orc.rh_build(rays, node="box_uv")
    # primary rays (rayhconstruct5DB.cl): d.x == 0 gives slope 0
    assert list(nodes[0][[0, 4, 1, 5, 2, 6]]) == [-1.0, 2.0, 0.1, 0.7, 0.0, 0.0]
    _, nodes = orc.rh_build(rays, use_valid=True, node="box_uv")
    # validity-aware (rayhconstructP5DB.cl): the d.x == 0 ray sets [-1, 1], dropping the slope 2
    # read before it; the later slopes 0.5 and -1 lie inside
    assert list(nodes[0][[0, 4, 1, 5]]) == [-1.0, 1.0, 0.1, 0.7]
    assert nodes[0][3] == 1
    # a flat ray first: later slopes widen [-1, 1]
    _, nodes = orc.rh_build(_row_rays([(0, 1, 0.3), (1, 3, 0.2)], 2, 1), use_valid=True,
                            node="box_uv")
    assert list(nodes[0][[0, 4]]) == [-1.0, 3.0]
    # an invalid flat ray is skipped: plain union of the valid slopes 2 and 0.5
    _, nodes = orc.rh_build(_row_rays(d[:3], 3, 1, valid=[1, 0, 1]), use_valid=True,
                            node="box_uv")
    assert list(nodes[0][[0, 4, 1, 5]]) == [0.5, 2.0, 0.1, 0.5]
    # no valid ray: the leaf is invalid
    _, nodes = orc.rh_build(_row_rays(d[:2], 2, 1, valid=[0, 0]), use_valid=True, node="box_uv")
    assert nodes[0][3] == 0


def test_rh_build_5db_row_major_order(orc):
    # 4x2 image in one tile: rows are read top to bottom, each left to right
    base = np.tile(np.float32([2, 1, 0.5]), (8, 1))      # slope 0.5
    a = base.copy()
    a[3] = (0, 1, 0.5)   # flat at row 0, column 3
    a[4] = (1, 3, 0.5)   # slope 3 at row 1, column 0: read after the flat ray
    _, nodes = orc.rh_build(_row_rays(a, 4, 2), use_valid=True, node="box_uv")
    assert list(nodes[0][[0, 4]]) == [-1.0, 3.0]
    b = base.copy()
    b[3] = (1, 3, 0.5)   # slope 3 first
    b[4] = (0, 1, 0.5)   # then the flat ray
    _, nodes = orc.rh_build(_row_rays(b, 4, 2), use_valid=True, node="box_uv")
    assert list(nodes[0][[0, 4]]) == [-1.0, 1.0]
    # primary rays merge: both orders give [0, 3]
    for r in (a, b):
        _, nodes = orc.rh_build(_row_rays(r, 4, 2), node="box_uv")
        assert list(nodes[0][[0, 4]]) == [0.0, 3.0]


def test_rh_build_5db_levels_merge(orc, scenes):
    # levelConstructP5DB.cl merges children by min/max only (no reset above the leaves)
    rays = scenes.random_rays(16 * 16, seed=70, width=16)
    rays["dx"][::7] = 0.0
    rays["valid"] = (np.random.default_rng(71).random(256) < 0.7).astype(np.uint8)
    levels, nodes = orc.rh_build(rays, use_valid=True, node="box_uv")
    (gw0, gh0), (gw1, gh1) = levels[0], levels[1]
    for j in range(gh1):
        for k in range(gw1):
            kids = [nodes[(2 * j + a) * gw0 + 2 * k + b] for a in (0, 1) for b in (0, 1)]
            kids = [c for c in kids if c[3] == 1]
            nd = nodes[gw0 * gh0 + j * gw1 + k]
            assert nd[3] == (1 if kids else 0)
            if kids:
                kk = np.array(kids)
                np.testing.assert_array_equal(nd[[0, 1, 8, 9, 10]], kk[:, [0, 1, 8, 9, 10]].min(0))
                np.testing.assert_array_equal(nd[[4, 5, 12, 13, 14]], kk[:, [4, 5, 12, 13, 14]].max(0))


def test_rh_5db_compares_raw_dz(orc):
    # The node keeps the rays' raw d.z while the cull normalises corner directions: a ray
    # toward a triangle below its origin hits it with a unit direction and misses it with the
    # same direction scaled by 5 (d.z 4 against corner z near 0.8).  Above the origin the y
    # rule passes every node and both hit.
    below = np.array([[-0.5, -3.2, 4, 0.5, -3.2, 4, 0, -2.8, 4]], np.float32)
    above = below * np.float32([1, -1, 1] * 3)
    p = orc.params("rayhierarchy", node="box_uv")
    for v9, dy in ((below, -1), (above, 1)):
        for scale in (1.0, 5.0):
            rays = _rays_of([(0, 0, 0)], [(0, dy * 0.6 * scale, 0.8 * scale)], 1, 1)
            got = orc.trace(v9, rays, p, orc.CLOSEST_PRIMARY)
            exp = -1 if (dy < 0 and scale == 5.0) else 0
            assert got["id"][0] == exp, (dy, scale)
            if exp == 0:
                assert abs(got["t"][0] * scale - 5.0) < 1e-5


def _above_scene(scenes, n, seed):
    """Triangles with every vertex y >= -1.4, above the camera at y = -1.5 and the origins of
    _below_rays."""
    return scenes.random_triangles(n, seed=seed, extent=1.0, size=0.4)


def _below_rays(scenes, seed, maxt):
    r = scenes.random_rays(256, seed=seed, lo=-1.5, hi=1.5, maxt=maxt, width=16)
    r["oy"] = np.random.default_rng(seed).uniform(-3.0, -1.5, 256).astype(np.float32)
    r["dy"] = np.abs(r["dy"])  # upward, toward the triangles
    r["valid"] = (np.arange(256) % 5 != 0).astype(np.uint8)
    return r


def test_rh_5db_all_pass_equals_brute_force(orc, scenes):
    # Every triangle lies above every ray origin, so some corner of each grown box is above
    # each node's origin centre and every 5DB cull passes: the hierarchy must return the
    # exhaustive strict nearest hit (primary, secondary) and occlusion.
    v9 = _above_scene(scenes, 128, 72)
    p = orc.params("rayhierarchy", node="box_uv")
    cam = scenes.pinhole_rays(16, 16, origin=(0.0, -1.5, -3.0), fov_deg=90.0)
    got = orc.trace(v9, cam, p, orc.CLOSEST_PRIMARY)
    ids, ts = _strict_brute(v9, cam)
    np.testing.assert_array_equal(got["id"], ids)
    np.testing.assert_array_equal(got["t"], ts)
    assert (ids >= 0).sum() > 20
    sec = _below_rays(scenes, 73, np.inf)
    got = orc.trace(v9, sec, p, orc.CLOSEST_SECONDARY)
    ids, _ = _strict_brute(v9, sec, valid_only=True)
    np.testing.assert_array_equal(got["id"], ids)
    assert (ids >= 0).sum() > 5
    sh = _below_rays(scenes, 74, 6.0)
    got = orc.trace(v9, sh, p, orc.OCCLUDED)
    exp = _strict_occ_brute(v9, sh)
    np.testing.assert_array_equal(got["occ"], exp)
    assert 5 < exp.sum() < 200


def test_raybvh_5db_all_pass_every_primitive(orc, scenes):
    # ray-bvh with box_uv: with every cull passing, primary AND secondary rays see every
    # primitive of a BVH leaf (intersection5DB_BVH2.cl:84 loops over all of them, unlike
    # intersection6DB_BVH2.cl:88), and occlusion is exhaustive.
    v9 = _paired_scene(40, seed=75)
    v9[:, 1::3] += np.float32(2.5)   # every vertex above y = 0.4 (a dyadic shift keeps the pairs)
    nodes, roots, order = orc.raybvh_handoff(v9, max_prims=4, bvh_height=3)
    assert (nodes[nodes["n_primitives"] > 0]["n_primitives"] >= 2).sum() >= 10
    p = orc.params("ray-bvh", node="box_uv")
    # rays from below aimed at the triangles' centroids
    g = np.random.default_rng(76)
    tgt = v9.reshape(-1, 3, 3).mean(1)[g.integers(0, len(v9), 256)]
    o = np.stack([g.uniform(-1.5, 1.5, 256), g.uniform(-3.0, -1.5, 256),
                  g.uniform(-1.5, 1.5, 256)], 1).astype(np.float32)
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    sec = _rays_of(o, d, 16, 16)
    sec["valid"] = (np.arange(256) % 5 != 0).astype(np.uint8)
    got = orc.trace_raybvh(v9, nodes, roots, order, sec, p, orc.CLOSEST_SECONDARY)
    ids, _ = _strict_brute(v9, sec, valid_only=True)
    np.testing.assert_array_equal(got["id"], ids)
    assert (ids >= 0).sum() > 100
    p6 = orc.params("ray-bvh", node="box_dir")
    got6 = orc.trace_raybvh(v9, nodes, roots, order, sec, p6, orc.CLOSEST_SECONDARY)
    assert (got6["id"] != ids).sum() > 0, "box_dir's first-primitive rule must differ here"
    cam = scenes.pinhole_rays(16, 16, origin=(0.0, -2.0, -3.0), fov_deg=90.0)
    got = orc.trace_raybvh(v9, nodes, roots, order, cam, p, orc.CLOSEST_PRIMARY)
    ids, ts = _strict_brute(v9, cam)
    np.testing.assert_array_equal(got["id"], ids)
    sh = _below_rays(scenes, 77, 8.0)
    got = orc.trace_raybvh(v9, nodes, roots, order, sh, p, orc.OCCLUDED)
    np.testing.assert_array_equal(got["occ"], _strict_occ_brute(v9, sh))


def test_raybvh_5db_skips_empty_nodes(orc):
    # intersection5DB_BVH.cl:18,56: a node with no primitives and primitive offset 0 is
    # skipped with its subtree.  Nodes 0 and 1 are leaves of one triangle each; node 2 is an
    # interior node over them (children 0 and 1), i.e. prim_offset 0 with no primitives.
    v9 = np.array([[-1, 1, 4, 1, 1, 4, 0, 3, 4], [-1, 1, 6, 1, 1, 6, 0, 3, 6]], np.float32)
    nodes = np.zeros(3, orc.GPU_NODE_DTYPE)
    for i, z in ((0, 4), (1, 6)):
        nodes[i] = (-1, 1, z, 1, 3, z, i, 1)
    nodes[2] = (-1, 1, 4, 1, 3, 6, 0, 0)
    order = np.arange(2, dtype=np.uint32)
    rays = _rays_of([(0, 0, 0)] * 4, [(0, 0.4, 1)] * 4, 2, 2)
    p = orc.params("ray-bvh", node="box_uv")
    got = orc.trace_raybvh(v9, nodes, np.array([2], np.uint32), order, rays, p, orc.CLOSEST_PRIMARY)
    assert np.all(got["id"] == -1)
    got = orc.trace_raybvh(v9, nodes, np.array([0, 1], np.uint32), order, rays, p,
                           orc.CLOSEST_PRIMARY)
    assert np.all(got["id"] == 0)
    occ = orc.trace_raybvh(v9, nodes, np.array([2], np.uint32), order, rays, p, orc.OCCLUDED)
    assert np.all(occ["occ"] == 0)


# ---------------------------------------------------------------- the per-ray path, restated
@pytest.mark.parametrize("accel", ["rayhierarchy", "ray-bvh"])
@pytest.mark.parametrize("node", ["box_dir", "ia", "box_uv"])
def test_search_trace_equals_reference_semantics(orc, scenes, accel, node):
    # The per-ray path walks a padded search BVH and applies the reference's cull chain only to
    # triangles that pass the strict leaf test.  Its answer must be the reference's: equal to
    # the depth-first traversal of the reference structure (without the repair pass, which the
    # library applies to both paths alike), for coherent and incoherent rays, every mode.
    v9 = scenes.cluster_scene(3000, seed=60)
    cam = scenes.pinhole_rays(32, 32, origin=(0.0, 0.0, -12.0), fov_deg=50.0)
    rnd = scenes.random_rays(32 * 32, seed=61, lo=-5.0, hi=5.0, maxt=np.inf, width=32)
    rnd["valid"] = (np.random.default_rng(62).random(32 * 32) < 0.8).astype(np.uint8)
    seg = scenes.random_rays(32 * 32, seed=63, lo=-5.0, hi=5.0, maxt=10.0, width=32)
    p = orc.params(accel, node=node, repair_run=0)
    handoff = orc.raybvh_handoff(v9) if accel == "ray-bvh" else None
    for mode, rays in ((0, cam), (1, rnd), (2, seg)):
        got = orc.search_trace(v9, rays, p, mode, handoff=handoff, nthreads=4)
        if accel == "ray-bvh":
            ref = orc.trace_raybvh(v9, *handoff, rays, p, mode)
        else:
            ref = orc.trace(v9, rays, p, mode)
        if mode == 2:
            np.testing.assert_array_equal(got["occ"], ref["occ"])
            assert 10 < ref["occ"].sum() < 1000
        else:
            np.testing.assert_array_equal(got["id"], ref["id"])
            np.testing.assert_array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
            assert (ref["id"] >= 0).sum() > 5
        st = got["stats"]
        # the search visits a small part of the tree per ray
        assert st.visits < 0.1 * st.rays * st.nodes


def test_search_trace_mask_and_counts(orc, scenes):
    v9 = scenes.cluster_scene(2000, seed=64)
    rays = scenes.random_rays(16 * 16, seed=65, lo=-5.0, hi=5.0, maxt=10.0, width=16)
    p = orc.params("rayhierarchy")
    full = orc.search_trace(v9, rays, p, orc.OCCLUDED)
    mask = (np.arange(256) % 2).astype(np.uint8)
    half = orc.search_trace(v9, rays, p, orc.OCCLUDED, mask=mask)
    np.testing.assert_array_equal(half["occ"], full["occ"] * mask)
    assert half["stats"].rays == 128
    assert 0 < half["stats"].visits < full["stats"].visits
    # one triangle per leaf: a full binary tree over the scene's triangles
    assert full["stats"].nodes == len(v9) - 1
