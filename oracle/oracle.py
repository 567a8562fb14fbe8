"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product library never loads it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

CLOSEST_PRIMARY, CLOSEST_SECONDARY, OCCLUDED = 0, 1, 2
_fp = ctypes.c_void_p


class OrcRays(ctypes.Structure):
    _fields_ = [("ox", _fp), ("oy", _fp), ("oz", _fp), ("dx", _fp), ("dy", _fp), ("dz", _fp),
                ("tmin", _fp), ("tmax", _fp), ("valid", _fp),
                ("W", ctypes.c_int32), ("H", ctypes.c_int32)]


class OrcParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("node", ctypes.c_int32), ("cx", ctypes.c_int32),
                ("cy", ctypes.c_int32), ("height", ctypes.c_int32),
                ("repair_run", ctypes.c_int32)]


class OrcStats(ctypes.Structure):
    _fields_ = [("cull_tests", ctypes.c_uint64), ("leaf_pairs", ctypes.c_uint64),
                ("ray_tests", ctypes.c_uint64), ("repair_rays", ctypes.c_uint64)]


GPU_NODE_DTYPE = np.dtype([("ax", "<f4"), ("ay", "<f4"), ("az", "<f4"), ("bx", "<f4"),
                           ("by", "<f4"), ("bz", "<f4"), ("prim_offset", "<u4"),
                           ("n_primitives", "<u4")])

_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        _LIB = ctypes.CDLL(str(LIB_PATH))
        _LIB.orc_trace.restype = ctypes.c_int
        _LIB.orc_trace_raybvh.restype = ctypes.c_int
        _LIB.orc_rh_build.restype = ctypes.c_int64
        _LIB.orc_bvh_build.restype = ctypes.c_int64
        _LIB.orc_raybvh_handoff.restype = ctypes.c_int64
        _LIB.orc_bvh_order.restype = ctypes.c_int64
        _LIB.orc_cpu_bvh_create.restype = ctypes.c_void_p
        _LIB.orc_cpu_bvh_destroy.argtypes = [ctypes.c_void_p]
        _LIB.orc_cpu_bvh_closest.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                             _fp, _fp, _fp, _fp]
        _LIB.orc_cpu_bvh_occluded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, _fp]
    return _LIB


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class _RayKeep:
    """Holds contiguous copies of a ray dict alive while the C struct points at them."""

    def __init__(self, r: dict):
        self.arrs = {k: _f32(r[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")}
        v = r.get("valid")
        self.valid = None if v is None else np.ascontiguousarray(v, dtype=np.uint8)
        a = self.arrs
        self.s = OrcRays(a["ox"].ctypes.data, a["oy"].ctypes.data, a["oz"].ctypes.data,
                         a["dx"].ctypes.data, a["dy"].ctypes.data, a["dz"].ctypes.data,
                         a["tmin"].ctypes.data, a["tmax"].ctypes.data,
                         None if self.valid is None else self.valid.ctypes.data,
                         int(r["width"]), int(r["height"]))
        self.n = int(r["width"]) * int(r["height"])


NODE_KINDS = {"box_dir": 0, "ia": 1, "box_uv": 2}


def params(kind="rayhierarchy", node="box_dir", cx=4, cy=4, height=3, repair_run=None):
    k = {"naive": 0, "rayhierarchy": 1, "ray-bvh": 2}[kind] if isinstance(kind, str) else kind
    nd = NODE_KINDS[node] if isinstance(node, str) else node
    if repair_run is None:
        repair_run = 1 if k == 1 else 0
    return OrcParams(k, nd, cx, cy, height, repair_run)


NEAR_TIE, ULP_MARGINAL = 1, 2  # bits of the per-ray flags


def trace(v9, rays: dict, p: OrcParams, mode: int, flags: bool = False):
    """Returns dict(id, t, b1, b2) for closest modes or dict(occ) for OCCLUDED, plus stats;
    with flags=True also "flags", one byte per ray (NEAR_TIE | ULP_MARGINAL bits)."""
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    keep = _RayKeep(rays)
    n = keep.n
    st = OrcStats()
    fl = np.zeros(n, np.uint8) if flags else None
    flp = _fp(fl.ctypes.data) if flags else None
    if mode == OCCLUDED:
        occ = np.zeros(n, np.uint8)
        rc = lib().orc_trace(_fp(v9.ctypes.data), ctypes.c_uint32(ntri), ctypes.byref(keep.s),
                             ctypes.byref(p), ctypes.c_int32(mode), None, None, None, None,
                             _fp(occ.ctypes.data), ctypes.byref(st), flp)
        assert rc == 0
        out = dict(occ=occ, stats=st)
    else:
        out = dict(id=np.zeros(n, np.int32), t=np.zeros(n, np.float32),
                   b1=np.zeros(n, np.float32), b2=np.zeros(n, np.float32))
        rc = lib().orc_trace(_fp(v9.ctypes.data), ctypes.c_uint32(ntri), ctypes.byref(keep.s),
                             ctypes.byref(p), ctypes.c_int32(mode), _fp(out["id"].ctypes.data),
                             _fp(out["t"].ctypes.data), _fp(out["b1"].ctypes.data),
                             _fp(out["b2"].ctypes.data), None, ctypes.byref(st), flp)
        assert rc == 0
        out["stats"] = st
    if flags:
        out["flags"] = fl
    return out


def trace_raybvh(v9, nodes, roots, prim_order, rays: dict, p: OrcParams, mode: int,
                 flags: bool = False):
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    nodes = np.ascontiguousarray(nodes, dtype=GPU_NODE_DTYPE)
    roots = np.ascontiguousarray(roots, dtype=np.uint32)
    prim_order = np.ascontiguousarray(prim_order, dtype=np.uint32)
    keep = _RayKeep(rays)
    n = keep.n
    st = OrcStats()
    args = [_fp(v9.ctypes.data), ctypes.c_uint32(ntri), _fp(nodes.ctypes.data),
            ctypes.c_uint32(nodes.size), _fp(roots.ctypes.data), ctypes.c_uint32(roots.size),
            _fp(prim_order.ctypes.data), ctypes.byref(keep.s), ctypes.byref(p), ctypes.c_int32(mode)]
    fl = np.zeros(n, np.uint8) if flags else None
    flp = _fp(fl.ctypes.data) if flags else None
    if mode == OCCLUDED:
        occ = np.zeros(n, np.uint8)
        rc = lib().orc_trace_raybvh(*args, None, None, None, None, _fp(occ.ctypes.data),
                                    ctypes.byref(st), flp)
        assert rc == 0
        out = dict(occ=occ, stats=st)
    else:
        out = dict(id=np.zeros(n, np.int32), t=np.zeros(n, np.float32),
                   b1=np.zeros(n, np.float32), b2=np.zeros(n, np.float32))
        rc = lib().orc_trace_raybvh(*args, _fp(out["id"].ctypes.data), _fp(out["t"].ctypes.data),
                                    _fp(out["b1"].ctypes.data), _fp(out["b2"].ctypes.data), None,
                                    ctypes.byref(st), flp)
        assert rc == 0
        out["stats"] = st
    if flags:
        out["flags"] = fl
    return out


def shading(v9, uv6, rays: dict, hit_id, naive=False):
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    keep = _RayKeep(rays)
    n = keep.n
    hit_id = np.ascontiguousarray(hit_id, dtype=np.int32)
    u, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
    dpdu, dpdv = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32)
    uvk = None if uv6 is None else _f32(uv6).reshape(-1)
    uvp = None if uvk is None else _fp(uvk.ctypes.data)
    lib().orc_shading(_fp(v9.ctypes.data), uvp, ctypes.c_uint32(ntri), ctypes.byref(keep.s),
                      _fp(hit_id.ctypes.data), 1 if naive else 0, _fp(u.ctypes.data),
                      _fp(v.ctypes.data),
                      _fp(dpdu.ctypes.data), _fp(dpdv.ctypes.data))
    return dict(u=u, v=v, dpdu=dpdu, dpdv=dpdv)


def rh_levels(W, H, cx=4, cy=4, height=3):
    dims = (ctypes.c_int32 * 128)()
    L = lib().orc_rh_levels(W, H, cx, cy, height, dims)
    return [(dims[2 * i], dims[2 * i + 1]) for i in range(L)]


def rh_well_formed(W, H, cx=4, cy=4, height=3) -> bool:
    return bool(lib().orc_rh_reference_well_formed(W, H, cx, cy, height))


def rh_build(rays: dict, cx=4, cy=4, height=3, use_valid=False, node="box_dir"):
    """Hierarchy nodes, 16 floats each: {dmin, valid, dmax, 0, omin, 0, omax, 0}; for node
    "box_uv" dmin/dmax hold (slope d.y/d.x, d.z, 0)."""
    levels = rh_levels(int(rays["width"]), int(rays["height"]), cx, cy, height)
    total = sum(w * h for w, h in levels)
    nodes = np.zeros((total, 16), np.float32)
    keep = _RayKeep(rays)
    nd = NODE_KINDS[node] if isinstance(node, str) else node
    got = lib().orc_rh_build(ctypes.byref(keep.s), cx, cy, height, 1 if use_valid else 0, nd,
                             _fp(nodes.ctypes.data))
    assert got == total
    return levels, nodes


def _vec(a):
    return (ctypes.c_float * len(a))(*[float(x) for x in a])


def mt(kind: str, o, d, v9):
    out = (ctypes.c_float * 3)()
    fn = {"strict": "orc_mt_strict", "naive": "orc_mt_naive", "naive_p": "orc_mt_naive_p",
          "repair": "orc_mt_repair", "pbrt": "orc_mt_pbrt"}[kind]
    ok = getattr(lib(), fn)(_vec(o), _vec(d), _vec(v9), out)
    return bool(ok), (out[0], out[1], out[2])


def cull(kind: str, node16, v9=None, bmin=None, bmax=None) -> bool:
    if kind == "tri":
        return bool(lib().orc_cull_tri(_vec(node16), _vec(v9)))
    fn = {"node6": lib().orc_cull_node6, "ia": lib().orc_cull_ia,
          "5db": lib().orc_cull_5db}[kind]
    return bool(fn(_vec(node16), _vec(bmin), _vec(bmax)))


SPLIT = {"sah": 0, "middle": 1, "equal": 2}


def bvh_build(v9, max_prims=4, gpu=True, bvh_height=3, split="sah"):
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    order = np.zeros(max(ntri, 1), np.uint32)
    nodes = np.zeros(max(2 * ntri, 1), GPU_NODE_DTYPE)
    top = ctypes.c_uint32()
    nn = lib().orc_bvh_build(_fp(v9.ctypes.data), ctypes.c_uint32(ntri), ctypes.c_uint32(max_prims),
                             ctypes.c_int32(SPLIT[split]), ctypes.c_int32(1 if gpu else 0),
                             ctypes.c_uint32(bvh_height), _fp(order.ctypes.data),
                             _fp(nodes.ctypes.data), ctypes.byref(top))
    assert nn >= 0
    return nodes[:nn].copy(), int(top.value), order[:ntri].copy()


def bvh_order(v9, max_prims=1, split="sah"):
    """BVHAccel's orderedPrims, repeats included (see orc_bvh_order)."""
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    cap = 16 * max(ntri, 1)
    order = np.zeros(cap, np.uint32)
    n = lib().orc_bvh_order(_fp(v9.ctypes.data), ctypes.c_uint32(ntri), ctypes.c_uint32(max_prims),
                            ctypes.c_int32(SPLIT[split]), _fp(order.ctypes.data),
                            ctypes.c_uint64(cap))
    assert n >= 0
    return order[:n].copy()


def raybvh_handoff(v9, max_prims=1, bvh_height=3, split="sah"):
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    nodes = np.zeros(max(2 * ntri + 8, 1), GPU_NODE_DTYPE)
    roots = np.zeros(max(2 * ntri + 8, 1), np.uint32)
    order = np.zeros(max(ntri, 1), np.uint32)
    nr = ctypes.c_uint32()
    nn = lib().orc_raybvh_handoff(_fp(v9.ctypes.data), ctypes.c_uint32(ntri),
                                  ctypes.c_uint32(max_prims), ctypes.c_int32(SPLIT[split]),
                                  ctypes.c_uint32(bvh_height),
                                  _fp(nodes.ctypes.data), _fp(roots.ctypes.data), ctypes.byref(nr),
                                  _fp(order.ctypes.data))
    return nodes[:nn].copy(), roots[:nr.value].copy(), order[:ntri].copy()


class CpuBvh:
    """pbrt's BVHAccel (SAH, maxnodeprims) with the inclusive Triangle predicate: the CPU
    baseline."""

    def __init__(self, v9, max_prims=4):
        self.v9 = _f32(v9).reshape(-1)
        self.h = lib().orc_cpu_bvh_create(_fp(self.v9.ctypes.data),
                                          ctypes.c_uint32(self.v9.size // 9),
                                          ctypes.c_uint32(max_prims))

    def closest(self, rays: dict, nthreads=0):
        keep = _RayKeep(rays)
        n = keep.n
        out = dict(id=np.zeros(n, np.int32), t=np.zeros(n, np.float32),
                   b1=np.zeros(n, np.float32), b2=np.zeros(n, np.float32))
        lib().orc_cpu_bvh_closest(self.h, ctypes.addressof(keep.s), nthreads,
                                  out["id"].ctypes.data, out["t"].ctypes.data,
                                  out["b1"].ctypes.data, out["b2"].ctypes.data)
        return out

    def occluded(self, rays: dict, nthreads=0):
        keep = _RayKeep(rays)
        occ = np.zeros(keep.n, np.uint8)
        lib().orc_cpu_bvh_occluded(self.h, ctypes.addressof(keep.s), nthreads, occ.ctypes.data)
        return occ

    def close(self):
        if self.h:
            lib().orc_cpu_bvh_destroy(self.h)
            self.h = None


def exact_closest(v9, rays: dict):
    v9 = _f32(v9).reshape(-1)
    keep = _RayKeep(rays)
    n = keep.n
    ids, ts = np.zeros(n, np.int32), np.zeros(n, np.float32)
    lib().orc_exact_closest(_fp(v9.ctypes.data), ctypes.c_uint32(v9.size // 9), ctypes.byref(keep.s),
                            _fp(ids.ctypes.data), _fp(ts.ctypes.data))
    return dict(id=ids, t=ts)


class OrcSearchStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("visits", ctypes.c_uint64),
                ("tests", ctypes.c_uint64), ("culls", ctypes.c_uint64),
                ("nodes", ctypes.c_uint64), ("chains", ctypes.c_uint64)]


def search_trace(v9, rays: dict, p: OrcParams, mode: int, handoff=None, mask=None, nthreads=0):
    """The library's per-ray path restated on the CPU (search BVH traversal + strict leaf test
    + exact cull chain).  handoff = (nodes, roots, prim_order) for ray-bvh.  Returns the trace
    outputs and stats (rays, visits, tests, culls, nodes)."""
    v9 = _f32(v9).reshape(-1)
    ntri = v9.size // 9
    keep = _RayKeep(rays)
    n = keep.n
    st = OrcSearchStats()
    if handoff is not None:
        nodes = np.ascontiguousarray(handoff[0], dtype=GPU_NODE_DTYPE)
        roots = np.ascontiguousarray(handoff[1], dtype=np.uint32)
        order = np.ascontiguousarray(handoff[2], dtype=np.uint32)
        hp = (_fp(nodes.ctypes.data), ctypes.c_uint32(nodes.size), _fp(roots.ctypes.data),
              ctypes.c_uint32(roots.size), _fp(order.ctypes.data))
    else:
        hp = (None, ctypes.c_uint32(0), None, ctypes.c_uint32(0), None)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = dict(id=np.zeros(n, np.int32), t=np.zeros(n, np.float32), b1=np.zeros(n, np.float32),
               b2=np.zeros(n, np.float32), occ=np.zeros(n, np.uint8))
    closest = mode != OCCLUDED
    rc = lib().orc_search_trace(
        _fp(v9.ctypes.data), ctypes.c_uint32(ntri), *hp, ctypes.byref(keep.s), ctypes.byref(p),
        ctypes.c_int32(mode), None if m is None else _fp(m.ctypes.data), ctypes.c_int32(nthreads),
        _fp(out["id"].ctypes.data) if closest else None,
        _fp(out["t"].ctypes.data) if closest else None,
        _fp(out["b1"].ctypes.data) if closest else None,
        _fp(out["b2"].ctypes.data) if closest else None,
        None if closest else _fp(out["occ"].ctypes.data), ctypes.byref(st))
    assert rc == 0
    if closest:
        del out["occ"]
    else:
        out = dict(occ=out["occ"])
    out["stats"] = st
    return out
