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


@pytest.mark.parametrize("split", ["sah", "middle", "equal"])
def test_handoff_matches_oracle_builder(rt, orc, scenes, split):
    # two independent restatements of BVHAccel + GPUbufferWide + parts agree exactly, for each
    # split method (bvh.cpp:214-221,332-427) and leaf size
    for v9 in (scenes.random_triangles(2500, seed=3), scenes.cluster_scene(61000, seed=4)):
        for mp in ((1, 4) if split == "sah" else (1, 2)):
            a = rt.raybvh_handoff(v9, mp, 3, split=split)
            b = orc.raybvh_handoff(v9, mp, 3, split=split)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    if split != "sah":
        # the split methods build different trees
        v9 = scenes.random_triangles(2500, seed=3)
        assert not np.array_equal(rt.raybvh_handoff(v9, 1, 3, split=split)[2],
                                  rt.raybvh_handoff(v9, 1, 3)[2])


def test_bvh_order_matches_oracle_builder(rt, orc, scenes):
    # the "sort" order of RayHieararchy: BVHAccel's primitive order, against the oracle's own
    # BVHAccel restatement, at leaf sizes 1 and 4 and over a scene with coincident centroids
    dup = np.repeat(scenes.random_triangles(200, seed=8), 3, axis=0)
    for v9 in (scenes.random_triangles(3000, seed=5), scenes.cluster_scene(20000, seed=6), dup):
        for split in ("sah", "middle", "equal"):
            for mp in (1, 4):
                a = rt.bvh_order(v9, mp, split=split)
                b = orc.bvh_build(v9, max_prims=mp, gpu=False, split=split)[2]
                assert np.array_equal(np.sort(a), np.arange(len(v9)))
                assert np.array_equal(a, b), (split, mp)
    assert rt.lib().rt_bvh_order(None, 3, 1, 0, None) == 1
    assert rt.lib().rt_bvh_order(None, 0, 1, 3, None) == 1  # unknown split method


def _first_occurrences(order):
    _, first = np.unique(order, return_index=True)
    return order[np.sort(first)]


def test_bvh_order_sah_leaf_fallthrough(rt, orc, scenes):
    # With more than 4 primitives per leaf the reference's SAH branch records a leaf and then
    # falls through to InitInterior (no return, bvh.cpp:417-428): orderedPrims lists those
    # primitives again.  The oracle keeps the repeats; rt_bvh_order returns the first
    # occurrences, the order that decides ties (a later copy never wins against the first).
    # groups of 7 copies of one triangle, shifted by 1e-4 along x (distinct centroids, boxes
    # of nearly the group's size): every split of a group costs about 7.125 > 7, so a group
    # takes the leaf branch once 7 <= max_prims; groups lie 10 apart and split cleanly
    base = np.array([0, 0, 0, 1, 1, 0, 1, 0, 1], np.float32)
    g = np.random.default_rng(9)
    tris = []
    for c in g.integers(-20, 20, size=(60, 3)) * 10.0:
        for i in g.permutation(7):
            t = base.reshape(3, 3) + c + np.array([i * 1e-4, 0, 0])
            tris.append(t.reshape(-1))
    v9 = np.array(tris, np.float32)
    for mp in (8, 16, 255):
        full = orc.bvh_order(v9, max_prims=mp)
        assert len(full) > len(v9), "the fall-through must occur here"
        a = rt.bvh_order(v9, mp)
        assert np.array_equal(a, _first_occurrences(full)), mp
    # at 4 and below there is no repeat
    assert len(orc.bvh_order(v9, max_prims=4)) == len(v9)
