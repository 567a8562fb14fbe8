"""GPU: the allocation-free trace (rt_trace_closest_noalloc) against rt_trace_closest.

Called through the C ABI.  Two calls in a row on each scene, the second with a smaller ray
image, must each give exactly the bits of rt_trace_closest (id, t, u, v, dpdu, dpdv):
rayhierarchy primary rays with the repair pass, ray-bvh secondary rays with validity, naive.
A ray image beyond the static scratch is refused with RT_ERR_TOO_LARGE.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(rt, sc, rays, mode, hits):
    rs, hs = rays.c_struct(), hits.c_struct()
    return rt.lib().rt_trace_closest_noalloc(sc.handle, ctypes.byref(rs), mode,
                                             ctypes.byref(hs), None)


def _np(h):
    return {k: getattr(h, k).cpu().numpy() for k in ("id", "t", "u", "v", "dpdu", "dpdv")}


def _same(a, b, label):
    for k in a:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        n = int((x != y).sum())
        assert n == 0, f"{label}: {n} {k} mismatches"


@pytest.mark.parametrize("accel", ["rayhierarchy", "ray-bvh", "naive"])
def test_noalloc_equals_trace_closest(rt, scenes, gpu_ctx, accel):
    import torch
    v9 = scenes.cluster_scene(8000, seed=70) if accel != "naive" else scenes.random_triangles(
        1024, seed=1)
    sc = rt.Scene(gpu_ctx, v9, desc=accel)
    mode = rt.RT_CLOSEST_SECONDARY if accel == "ray-bvh" else rt.RT_CLOSEST_PRIMARY
    big = scenes.pinhole_rays(160, 96, origin=(0.0, 0.0, -12.0), fov_deg=50.0)
    small = scenes.random_rays(48 * 32, seed=71, lo=-5.0, hi=5.0, maxt=np.inf, width=48)
    small["valid"] = (np.random.default_rng(72).random(48 * 32) < 0.8).astype(np.uint8)
    for k, r in enumerate((big, small)):
        print(f"[noalloc] {accel} call {k}: {r['width']}x{r['height']}", flush=True)
        d = rt.DeviceRays.from_numpy(r)
        ref = _np(sc.trace_closest(d, mode, shading=True))
        h = rt.DeviceHits.empty(d.count, shading=True)
        st = _noalloc(rt, sc, d, mode, h)
        torch.cuda.synchronize()
        assert st == 0, rt.lib().rt_error_string(st)
        _same(_np(h), ref, f"{accel} call {k}")
        if k == 0:
            assert (ref["id"] >= 0).sum() > 100
    sc.close()


def test_noalloc_refuses_too_large(rt, scenes, gpu_ctx):
    import torch
    sc = rt.Scene(gpu_ctx, scenes.random_triangles(256, seed=3), desc="rayhierarchy")
    W, H = 2048, 2048  # 4M rays: more than the static scratch holds
    z = torch.zeros(W * H, dtype=torch.float32, device="cuda")
    d = rt.DeviceRays(z, z, z, z, z, torch.ones_like(z), z, z, None, W, H)
    h = rt.DeviceHits.empty(W * H)
    assert _noalloc(rt, sc, d, rt.RT_CLOSEST_PRIMARY, h) == 7  # RT_ERR_TOO_LARGE
    sc.close()
