"""The C-ABI library loads and exports every function declared in include/*.h (no GPU
calls: only loading, symbol lookup and the host-only entry points)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(rt):
    L = rt.lib()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert set(names) <= set(rt.EXPORTS)


def test_error_strings_and_defaults(rt):
    L = rt.lib()
    assert L.rt_error_string(0) == b"ok"
    assert L.rt_error_string(4) == b"unsupported parameter combination"
    d = rt.default_desc("rayhierarchy")
    assert (d.chunk_x, d.chunk_y, d.height, d.repair_run, d.node) == (4, 4, 3, 1, 0)
    d = rt.default_desc("ray-bvh")
    assert (d.bvh_height, d.max_bvh_prim, d.repair_run) == (3, 1, 0)
    with pytest.raises(rt.RtError):
        rt.default_desc(7)


def test_invalid_arguments_rejected_without_device(rt):
    L = rt.lib()
    assert L.rt_ctx_create(0, None) == 1
    assert L.rt_trace_closest(None, None, 0, None, None) == 1
    assert L.rt_trace_occluded(None, None, None, None) == 1


def test_handoff_matches_oracle_builder(rt, orc, scenes):
    # two independent restatements of BVHAccel + GPUbufferWide + parts agree exactly
    for v9 in (scenes.random_triangles(2500, seed=3), scenes.cluster_scene(61000, seed=4)):
        a = rt.raybvh_handoff(v9, 1, 3)
        b = orc.raybvh_handoff(v9, 1, 3)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_bvh_order_matches_oracle_builder(rt, orc, scenes):
    # the "sort" order of RayHieararchy: BVHAccel's primitive order, against the oracle's own
    # BVHAccel restatement, at leaf sizes 1 and 4 and over a scene with coincident centroids
    dup = np.repeat(scenes.random_triangles(200, seed=8), 3, axis=0)
    for v9 in (scenes.random_triangles(3000, seed=5), scenes.cluster_scene(20000, seed=6), dup):
        for mp in (1, 4):
            a = rt.bvh_order(v9, mp)
            b = orc.bvh_build(v9, max_prims=mp, gpu=False)[2]
            assert np.array_equal(np.sort(a), np.arange(len(v9)))
            assert np.array_equal(a, b)
    assert rt.lib().rt_bvh_order(None, 3, 1, None) == 1
