"""pbrt-v2_amd -- MI355X (gfx950) ray-space accelerators: naive, rayhierarchy, ray-bvh.

Python view of the C ABI in include/pbrt_amd_rt.h (librt_amd.so, built by csrc/Makefile).
torch is used only for device memory and streams; every computation runs in the HIP
library.  Importing this package never falls back to a CPU path: a missing library raises.

The directory name is not a Python identifier; load it with ``load_package()`` from the
repository root helpers (tests/conftest.py, bench.py, __graft_entry__.py) or importlib.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "librt_amd.so"

RT_OK = 0
RT_ACCEL_NAIVE, RT_ACCEL_RAYHIERARCHY, RT_ACCEL_RAYBVH = 0, 1, 2
RT_NODE_BOX_DIR, RT_NODE_IA, RT_NODE_BOX_UV = 0, 1, 2
NODE_NAMES = {"box_dir": RT_NODE_BOX_DIR, "ia": RT_NODE_IA, "box_uv": RT_NODE_BOX_UV}
RT_CLOSEST_PRIMARY, RT_CLOSEST_SECONDARY = 0, 1
FLAG_BVH_STACK, FLAG_PAIR_CAP = 1, 2  # rt_ctx_check bits (RT_FLAG_*)

ACCEL_NAMES = {"naive": RT_ACCEL_NAIVE, "rayhierarchy": RT_ACCEL_RAYHIERARCHY,
               "ray-bvh": RT_ACCEL_RAYBVH}

# exported C symbols (every function declared in include/pbrt_amd_rt.h)
EXPORTS = (
    "rt_accel_desc_default", "rt_ctx_create", "rt_ctx_destroy", "rt_ctx_check", "rt_scene_create",
    "rt_scene_set_bvh", "rt_scene_destroy", "rt_trace_closest", "rt_trace_occluded",
    "rt_make_shadow_rays", "rt_scene_set_timing", "rt_scene_kernel_time",
    "rt_error_string", "rt_raybvh_build_handoff", "rt_raybvh_free_handoff", "rt_bvh_order", "rt_scene_set_param",
    "rt_batch_create", "rt_batch_destroy", "rt_batch_clear", "rt_batch_add_closest",
    "rt_batch_add_occluded", "rt_batch_run", "rt_batch_set_param", "rt_batch_get_param",
    "rt_batch_autotune", "rt_batch_set_timing", "rt_batch_kernel_time",
    "rt_server_create", "rt_server_submit", "rt_server_wait", "rt_server_counters",
    "rt_server_destroy", "rt_make_bounce_rays", "rt_trace_closest_noalloc",
)
(RT_PARAM_CLUSTERS_CLOSEST, RT_PARAM_CLUSTERS_OCCLUDED, RT_PARAM_GRID_CLOSEST,
 RT_PARAM_GRID_OCCLUDED, RT_PARAM_GRAPH, RT_PARAM_RAY_CLUSTERS_CLOSEST,
 RT_PARAM_RAY_CLUSTERS_OCCLUDED, RT_PARAM_RAY_GRID_CLOSEST,
 RT_PARAM_RAY_GRID_OCCLUDED, RT_PARAM_WALK_CLOSEST, RT_PARAM_WALK_OCCLUDED,
 RT_PARAM_TILE_CLOSEST, RT_PARAM_TILE_OCCLUDED) = range(1, 14)
# a per-ray threshold above any cluster count: every block enumerates its candidate pairs
RAY_PATH_OFF = 1 << 30
# kernel_time modes: + RT_TIME_PER_RAY = only the per-ray kernel, + RT_TIME_UNIT = only the
# unit (block) kernel of the launches of a mode
RT_TIME_PER_RAY, RT_TIME_UNIT = 3, 6


class RtError(RuntimeError):
    def __init__(self, status: int, what: str, msg: str):
        super().__init__(f"{what} failed: status {status} ({msg})")
        self.status = status


class AccelDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("node", ctypes.c_int32),
                ("chunk_x", ctypes.c_int32), ("chunk_y", ctypes.c_int32),
                ("height", ctypes.c_int32), ("repair_run", ctypes.c_int32),
                ("bvh_height", ctypes.c_int32), ("max_bvh_prim", ctypes.c_int32)]


class GpuNode(ctypes.Structure):
    _fields_ = [("ax", ctypes.c_float), ("ay", ctypes.c_float), ("az", ctypes.c_float),
                ("bx", ctypes.c_float), ("by", ctypes.c_float), ("bz", ctypes.c_float),
                ("prim_offset", ctypes.c_uint32), ("n_primitives", ctypes.c_uint32)]


GPU_NODE_DTYPE = np.dtype([("ax", "<f4"), ("ay", "<f4"), ("az", "<f4"), ("bx", "<f4"),
                           ("by", "<f4"), ("bz", "<f4"), ("prim_offset", "<u4"),
                           ("n_primitives", "<u4")])

_fp = ctypes.c_void_p


class RtRequest(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("spp", ctypes.c_int32), ("origin", ctypes.c_float * 3),
                ("fov_deg", ctypes.c_float), ("light", ctypes.c_float * 3),
                ("seed", ctypes.c_uint32), ("host_id", ctypes.c_void_p),
                ("host_t", ctypes.c_void_p), ("host_occluded", ctypes.c_void_p)]


class RtRequestStats(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("primary_rays", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("shadow_rays", ctypes.c_uint64)]


class RtRays(ctypes.Structure):
    _fields_ = [("ox", _fp), ("oy", _fp), ("oz", _fp), ("dx", _fp), ("dy", _fp), ("dz", _fp),
                ("tmin", _fp), ("tmax", _fp), ("valid", _fp),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class RtHits(ctypes.Structure):
    _fields_ = [("id", _fp), ("t", _fp), ("u", _fp), ("v", _fp), ("dpdu", _fp), ("dpdv", _fp)]


_LIB = None


def lib() -> ctypes.CDLL:
    """Loads librt_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR / 'csrc'}` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    L.rt_accel_desc_default.argtypes = [ctypes.c_int32, P(AccelDesc)]
    L.rt_ctx_create.argtypes = [ctypes.c_int32, P(ctypes.c_void_p)]
    L.rt_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.rt_ctx_check.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32)]
    L.rt_scene_create.argtypes = [ctypes.c_void_p, _fp, _fp, ctypes.c_uint32, P(AccelDesc),
                                  P(ctypes.c_void_p)]
    L.rt_scene_set_bvh.argtypes = [ctypes.c_void_p, _fp, ctypes.c_uint32, _fp, ctypes.c_uint32, _fp]
    L.rt_scene_destroy.argtypes = [ctypes.c_void_p]
    L.rt_trace_closest.argtypes = [ctypes.c_void_p, P(RtRays), ctypes.c_int32, P(RtHits), _fp]
    L.rt_trace_occluded.argtypes = [ctypes.c_void_p, P(RtRays), _fp, _fp]
    L.rt_make_shadow_rays.argtypes = [P(RtRays), _fp, _fp, P(ctypes.c_float)] + [_fp] * 9 + [_fp]
    L.rt_scene_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.rt_scene_set_param.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.rt_scene_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_float),
                                       P(ctypes.c_int32)]
    L.rt_error_string.argtypes = [ctypes.c_int]
    L.rt_error_string.restype = ctypes.c_char_p
    L.rt_raybvh_build_handoff.argtypes = [_fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                          ctypes.c_uint32, P(ctypes.c_void_p), P(ctypes.c_uint32),
                                          P(ctypes.c_void_p), P(ctypes.c_uint32),
                                          P(ctypes.c_void_p)]
    L.rt_raybvh_free_handoff.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.rt_bvh_order.argtypes = [_fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _fp]
    L.rt_batch_create.argtypes = [ctypes.c_void_p, P(ctypes.c_void_p)]
    L.rt_batch_destroy.argtypes = [ctypes.c_void_p]
    L.rt_batch_clear.argtypes = [ctypes.c_void_p]
    L.rt_batch_add_closest.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(RtRays), ctypes.c_int32,
                                       P(RtHits)]
    L.rt_batch_add_occluded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(RtRays), _fp]
    L.rt_batch_run.argtypes = [ctypes.c_void_p, _fp]
    L.rt_batch_set_param.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.rt_batch_get_param.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_int32)]
    L.rt_batch_autotune.argtypes = [ctypes.c_void_p, _fp, ctypes.c_int32]
    L.rt_batch_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.rt_batch_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_float),
                                       P(ctypes.c_int32)]
    L.rt_server_create.argtypes = [ctypes.c_void_p, P(ctypes.c_void_p), ctypes.c_int32,
                                   ctypes.c_uint64, ctypes.c_int32, P(ctypes.c_void_p)]
    L.rt_server_submit.argtypes = [ctypes.c_void_p, P(RtRequest), P(ctypes.c_uint64)]
    L.rt_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double,
                                 P(RtRequestStats)]
    L.rt_server_counters.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_int32)]
    L.rt_server_destroy.argtypes = [ctypes.c_void_p]
    L.rt_make_bounce_rays.argtypes = [ctypes.c_void_p, P(RtRays), _fp, _fp, ctypes.c_uint32,
                                      ctypes.c_int32] + [_fp] * 9 + [_fp]
    _LIB = L
    return L


def check(status: int, what: str) -> None:
    if status != RT_OK:
        msg = lib().rt_error_string(status).decode()
        raise RtError(status, what, msg)


def default_desc(kind, **overrides) -> AccelDesc:
    """Reference defaults of the named accelerator (naive / rayhierarchy / ray-bvh)."""
    if isinstance(kind, str):
        kind = ACCEL_NAMES[kind]
    d = AccelDesc()
    check(lib().rt_accel_desc_default(kind, ctypes.byref(d)), "rt_accel_desc_default")
    for k, v in overrides.items():
        if k == "node" and isinstance(v, str):
            v = NODE_NAMES[v]
        setattr(d, k, int(v))
    return d


SPLIT_METHODS = {"sah": 0, "middle": 1, "equal": 2}


def raybvh_handoff(v9: np.ndarray, max_prims: int = 1, bvh_height: int = 3, split="sah"):
    """CPU-built ray-bvh handoff (global BVH order, 60k-triangle parts, per-part BVH with the
    named split method, breadth-first GPUNode layout).  Returns (nodes, roots, prim_order)."""
    L = lib()
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1)
    ntri = v9.size // 9
    nodes_p, roots_p, order_p = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nn, nr = ctypes.c_uint32(), ctypes.c_uint32()
    check(L.rt_raybvh_build_handoff(v9.ctypes.data, ntri, max_prims, SPLIT_METHODS[split],
                                    bvh_height,
                                    ctypes.byref(nodes_p), ctypes.byref(nn), ctypes.byref(roots_p),
                                    ctypes.byref(nr), ctypes.byref(order_p)),
          "rt_raybvh_build_handoff")
    try:
        nodes = np.ctypeslib.as_array(ctypes.cast(nodes_p, ctypes.POINTER(ctypes.c_uint8)),
                                      shape=(nn.value * 32,)).view(GPU_NODE_DTYPE).copy()
        roots = np.ctypeslib.as_array(ctypes.cast(roots_p, ctypes.POINTER(ctypes.c_uint32)),
                                      shape=(max(nr.value, 1),))[:nr.value].copy()
        order = np.ctypeslib.as_array(ctypes.cast(order_p, ctypes.POINTER(ctypes.c_uint32)),
                                      shape=(max(ntri, 1),))[:ntri].copy()
    finally:
        L.rt_raybvh_free_handoff(nodes_p, roots_p, order_p)
    return nodes, roots, order


def bvh_order(v9: np.ndarray, max_prims: int = 1, split="sah") -> np.ndarray:
    """BVHAccel primitive order over the triangles (position -> input index), the order
    RayHieararchy's "sort" parameter swaps its triangles into."""
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1)
    ntri = v9.size // 9
    order = np.zeros(max(ntri, 1), np.uint32)
    check(lib().rt_bvh_order(v9.ctypes.data, ntri, max_prims, SPLIT_METHODS[split],
                             order.ctypes.data), "rt_bvh_order")
    return order[:ntri].copy()


def _torch():
    import torch
    return torch


@dataclass
class DeviceRays:
    """A ray image in device memory (torch tensors on the HIP device)."""
    ox: object
    oy: object
    oz: object
    dx: object
    dy: object
    dz: object
    tmin: object
    tmax: object
    valid: object  # uint8 tensor or None
    width: int
    height: int

    @property
    def count(self) -> int:
        return self.width * self.height

    @staticmethod
    def from_numpy(r: dict, device="cuda"):
        torch = _torch()

        def f(a, dt=torch.float32):
            return torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt)

        valid = r.get("valid")
        return DeviceRays(f(r["ox"]), f(r["oy"]), f(r["oz"]), f(r["dx"]), f(r["dy"]), f(r["dz"]),
                          f(r["tmin"]), f(r["tmax"]),
                          None if valid is None else f(valid, torch.uint8),
                          int(r["width"]), int(r["height"]))

    def c_struct(self) -> RtRays:
        p = lambda t: None if t is None else t.data_ptr()
        return RtRays(p(self.ox), p(self.oy), p(self.oz), p(self.dx), p(self.dy), p(self.dz),
                      p(self.tmin), p(self.tmax), p(self.valid), self.width, self.height)


@dataclass
class DeviceHits:
    id: object
    t: object
    u: object
    v: object
    dpdu: object = None
    dpdv: object = None

    @staticmethod
    def empty(n: int, device="cuda", shading: bool = False):
        torch = _torch()
        h = DeviceHits(torch.empty(n, dtype=torch.int32, device=device),
                       torch.empty(n, dtype=torch.float32, device=device),
                       torch.empty(n, dtype=torch.float32, device=device),
                       torch.empty(n, dtype=torch.float32, device=device))
        if shading:
            h.dpdu = torch.empty(3 * n, dtype=torch.float32, device=device)
            h.dpdv = torch.empty(3 * n, dtype=torch.float32, device=device)
        return h

    def c_struct(self) -> RtHits:
        p = lambda t: None if t is None else t.data_ptr()
        return RtHits(p(self.id), p(self.t), p(self.u), p(self.v), p(self.dpdu), p(self.dpdv))


def _stream_ptr(stream):
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


class Context:
    def __init__(self, device: int = 0):
        self.handle = ctypes.c_void_p()
        check(lib().rt_ctx_create(device, ctypes.byref(self.handle)), "rt_ctx_create")
        self.device = device

    def check(self) -> int:
        """RT_FLAG_* bits the traversal kernels recorded since the last check (then cleared):
        FLAG_BVH_STACK, a ray-bvh block's traversal stack would have overflowed; FLAG_PAIR_CAP,
        a block's pair queue was full.  0 means every trace completed its traversal."""
        f = ctypes.c_uint32(0)
        check(lib().rt_ctx_check(self.handle, ctypes.byref(f)), "rt_ctx_check")
        return f.value

    def close(self):
        if self.handle:
            lib().rt_ctx_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class Scene:
    """A resident scene behind one accelerator (the plain per-scene form)."""

    def __init__(self, ctx: Context, v9: np.ndarray, uv6: np.ndarray | None = None,
                 desc: AccelDesc | str = "rayhierarchy", build_bvh: bool = True):
        if isinstance(desc, str):
            desc = default_desc(desc)
        self.desc = desc
        self.ctx = ctx
        self.v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1)
        self.ntri = self.v9.size // 9
        self.uv6 = None if uv6 is None else np.ascontiguousarray(uv6, dtype=np.float32).reshape(-1)
        self.handle = ctypes.c_void_p()
        check(lib().rt_scene_create(ctx.handle, self.v9.ctypes.data,
                                    None if self.uv6 is None else self.uv6.ctypes.data,
                                    self.ntri, ctypes.byref(desc), ctypes.byref(self.handle)),
              "rt_scene_create")
        self.handoff = None
        if desc.kind == RT_ACCEL_RAYBVH and build_bvh and self.ntri:
            self.set_bvh(*raybvh_handoff(self.v9, desc.max_bvh_prim, desc.bvh_height))

    def set_bvh(self, nodes: np.ndarray, roots: np.ndarray, prim_order: np.ndarray):
        nodes = np.ascontiguousarray(nodes, dtype=GPU_NODE_DTYPE)
        roots = np.ascontiguousarray(roots, dtype=np.uint32)
        prim_order = np.ascontiguousarray(prim_order, dtype=np.uint32)
        check(lib().rt_scene_set_bvh(self.handle, nodes.ctypes.data, nodes.size,
                                     roots.ctypes.data, roots.size, prim_order.ctypes.data),
              "rt_scene_set_bvh")
        self.handoff = (nodes, roots, prim_order)

    def trace_closest(self, rays: DeviceRays, mode: int = RT_CLOSEST_PRIMARY, hits=None,
                      stream=None, shading: bool = False) -> DeviceHits:
        if hits is None:
            hits = DeviceHits.empty(rays.count, device=rays.ox.device, shading=shading)
        rs, hs = rays.c_struct(), hits.c_struct()
        check(lib().rt_trace_closest(self.handle, ctypes.byref(rs), mode, ctypes.byref(hs),
                                     _stream_ptr(stream)), "rt_trace_closest")
        return hits

    def trace_occluded(self, rays: DeviceRays, out=None, stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty(rays.count, dtype=torch.uint8, device=rays.ox.device)
        rs = rays.c_struct()
        check(lib().rt_trace_occluded(self.handle, ctypes.byref(rs), out.data_ptr(),
                                      _stream_ptr(stream)), "rt_trace_occluded")
        return out

    def set_param(self, param: int, value: int):
        check(lib().rt_scene_set_param(self.handle, param, int(value)), "rt_scene_set_param")

    def set_timing(self, on: bool):
        check(lib().rt_scene_set_timing(self.handle, 1 if on else 0), "rt_scene_set_timing")

    def kernel_time(self, mode: int):
        """(total ms, launches) of the traversal kernels of one mode recorded since
        set_timing(True); mode 0/1 closest primary/secondary, 2 occlusion (+ RT_TIME_PER_RAY:
        the per-ray kernel only, + RT_TIME_UNIT: the unit kernel only)."""
        ms, n = ctypes.c_float(), ctypes.c_int32()
        check(lib().rt_scene_kernel_time(self.handle, mode, ctypes.byref(ms), ctypes.byref(n)),
              "rt_scene_kernel_time")
        return ms.value, n.value

    def close(self):
        if self.handle:
            lib().rt_scene_destroy(self.handle)
            self.handle = ctypes.c_void_p()


def make_bounce_rays(scene: "Scene", rays: DeviceRays, hits: DeviceHits, seed: int = 0,
                     depth: int = 0, stream=None, out: DeviceRays | None = None) -> DeviceRays:
    """Cosine-weighted path bounce rays from closest hits, built on the device."""
    torch = _torch()
    n = rays.count
    dev = rays.ox.device
    if out is None:
        f = lambda: torch.empty(n, dtype=torch.float32, device=dev)
        out = DeviceRays(f(), f(), f(), f(), f(), f(), f(), f(),
                         torch.empty(n, dtype=torch.uint8, device=dev), rays.width, rays.height)
    rs = rays.c_struct()
    check(lib().rt_make_bounce_rays(scene.handle, ctypes.byref(rs), hits.id.data_ptr(),
                                    hits.t.data_ptr(), seed, depth, out.ox.data_ptr(),
                                    out.oy.data_ptr(), out.oz.data_ptr(), out.dx.data_ptr(),
                                    out.dy.data_ptr(), out.dz.data_ptr(), out.tmin.data_ptr(),
                                    out.tmax.data_ptr(), out.valid.data_ptr(),
                                    _stream_ptr(stream)), "rt_make_bounce_rays")
    return out


class Batch:
    """Batched form: many rayhierarchy or ray-bvh jobs (scenes x ray images, closest and
    occlusion) run as one launch sequence per (mode, node type, accelerator) group with a
    shared persistent traversal kernel.
    The jobs' tensors are kept referenced until the batch is cleared."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.handle = ctypes.c_void_p()
        check(lib().rt_batch_create(ctx.handle, ctypes.byref(self.handle)), "rt_batch_create")
        self._keep = []

    def add_closest(self, scene: Scene, rays: DeviceRays, mode: int = RT_CLOSEST_PRIMARY,
                    hits: DeviceHits | None = None, shading: bool = False) -> DeviceHits:
        if hits is None:
            hits = DeviceHits.empty(rays.count, device=rays.ox.device, shading=shading)
        rs, hs = rays.c_struct(), hits.c_struct()
        check(lib().rt_batch_add_closest(self.handle, scene.handle, ctypes.byref(rs), mode,
                                         ctypes.byref(hs)), "rt_batch_add_closest")
        self._keep.append((scene, rays, hits))
        return hits

    def add_occluded(self, scene: Scene, rays: DeviceRays, out=None):
        torch = _torch()
        if out is None:
            out = torch.empty(rays.count, dtype=torch.uint8, device=rays.ox.device)
        rs = rays.c_struct()
        check(lib().rt_batch_add_occluded(self.handle, scene.handle, ctypes.byref(rs),
                                          out.data_ptr()), "rt_batch_add_occluded")
        self._keep.append((scene, rays, out))
        return out

    def run(self, stream=None):
        check(lib().rt_batch_run(self.handle, _stream_ptr(stream)), "rt_batch_run")

    def clear(self):
        check(lib().rt_batch_clear(self.handle), "rt_batch_clear")
        self._keep = []

    def set_param(self, param: int, value: int):
        check(lib().rt_batch_set_param(self.handle, param, int(value)), "rt_batch_set_param")

    def get_param(self, param: int) -> int:
        v = ctypes.c_int32()
        check(lib().rt_batch_get_param(self.handle, param, ctypes.byref(v)), "rt_batch_get_param")
        return v.value

    def autotune(self, reps: int = 3, stream=None):
        check(lib().rt_batch_autotune(self.handle, _stream_ptr(stream), reps), "rt_batch_autotune")

    def set_timing(self, on: bool):
        check(lib().rt_batch_set_timing(self.handle, 1 if on else 0), "rt_batch_set_timing")

    def kernel_time(self, mode: int):
        ms, n = ctypes.c_float(), ctypes.c_int32()
        check(lib().rt_batch_kernel_time(self.handle, mode, ctypes.byref(ms), ctypes.byref(n)),
              "rt_batch_kernel_time")
        return ms.value, n.value

    def close(self):
        if self.handle:
            lib().rt_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()
        self._keep = []


class Server:
    """Serving form: a request loop over resident rayhierarchy / ray-bvh scenes with continuous
    batching.  submit() returns a ticket at once; wait() returns the request's stats and, if
    outputs were asked for, numpy arrays of hit ids, t and occlusion flags."""

    def __init__(self, ctx: Context, scenes: list, mem_budget: int = 0, max_active: int = 64):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (ctypes.c_void_p * len(self.scenes))(*[s.handle.value for s in self.scenes])
        self.handle = ctypes.c_void_p()
        check(lib().rt_server_create(ctx.handle, arr, len(self.scenes), mem_budget, max_active,
                                     ctypes.byref(self.handle)), "rt_server_create")
        self._out = {}

    def submit(self, scene: int, width: int, height: int, spp: int = 1,
               origin=(0.0, 0.0, -3.0), fov_deg: float = 60.0, light=(2.0, 2.0, -2.0),
               seed: int = 0, outputs: bool = False) -> int:
        n = width * spp * height
        out = None
        q = RtRequest(scene, width, height, spp, (ctypes.c_float * 3)(*origin), fov_deg,
                      (ctypes.c_float * 3)(*light), seed, None, None, None)
        if outputs:
            out = (np.empty(n, np.int32), np.empty(n, np.float32), np.empty(n, np.uint8))
            q.host_id, q.host_t, q.host_occluded = (a.ctypes.data for a in out)
        t = ctypes.c_uint64()
        check(lib().rt_server_submit(self.handle, ctypes.byref(q), ctypes.byref(t)),
              "rt_server_submit")
        self._out[t.value] = out
        return t.value

    def wait(self, ticket: int, timeout_s: float = -1.0):
        st = RtRequestStats()
        check(lib().rt_server_wait(self.handle, ticket, timeout_s, ctypes.byref(st)),
              "rt_server_wait")
        out = self._out.pop(ticket, None)
        stats = dict(submit_s=st.submit_s, admit_s=st.admit_s, done_s=st.done_s,
                     latency_s=st.done_s - st.submit_s, primary_rays=st.primary_rays,
                     shadow_rays=st.shadow_rays)
        if out is not None:
            stats.update(id=out[0], t=out[1], occ=out[2])
        return stats

    def counters(self):
        t, m = ctypes.c_uint64(), ctypes.c_int32()
        check(lib().rt_server_counters(self.handle, ctypes.byref(t), ctypes.byref(m)),
              "rt_server_counters")
        return t.value, m.value

    def close(self):
        if self.handle:
            lib().rt_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()


def make_shadow_rays(primary: DeviceRays, hits: DeviceHits, light, stream=None,
                     out: DeviceRays | None = None) -> DeviceRays:
    """Point-light shadow rays from closest hits, built on the device (into `out` if given)."""
    torch = _torch()
    n = primary.count
    dev = primary.ox.device
    if out is None:
        f = lambda: torch.empty(n, dtype=torch.float32, device=dev)
        out = DeviceRays(f(), f(), f(), f(), f(), f(), f(), f(),
                         torch.empty(n, dtype=torch.uint8, device=dev), primary.width,
                         primary.height)
    ps = primary.c_struct()
    lt = (ctypes.c_float * 3)(*[float(x) for x in light])
    check(lib().rt_make_shadow_rays(ctypes.byref(ps), hits.id.data_ptr(), hits.t.data_ptr(), lt,
                                    out.ox.data_ptr(), out.oy.data_ptr(), out.oz.data_ptr(),
                                    out.dx.data_ptr(), out.dy.data_ptr(), out.dz.data_ptr(),
                                    out.tmin.data_ptr(), out.tmax.data_ptr(),
                                    out.valid.data_ptr(), _stream_ptr(stream)),
          "rt_make_shadow_rays")
    return out


__all__ = [n for n in dir() if not n.startswith("_")]
del os
