"""Seeded synthetic scenes and ray images of the benchmark workloads (numpy, FP32).

All arithmetic that defines a ray is done in float32 so the same arrays can be handed to the
device and to a CPU checker.  Workloads:
  random1k     1,024 random triangles (centroid U[-1,1]^3, vertex offsets U[-0.1,0.1]^3),
               128x128 pinhole camera at (0,0,-3), 60 deg vertical FOV, light (2,2,-2)
  torus        UV torus R=1, r=0.4, 128x96 quads (24,576 tris) + 2-triangle ground plane +
               200 small random triangles; 512x512 camera
  room         closed 10x6x20 box room (12 tris) + 400 Gaussian clusters of 200 triangles
  clusters     Gaussian triangle clusters in [-5,5]^3 (the batched/serving scene generator)
  uneven_batch 64 cluster scenes of log-uniform size with camera + random shadow rays
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
INF = np.float32(np.inf)


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# ---------------------------------------------------------------------------------------
# scenes (v9: float32 [ntri, 9])

def random_triangles(n: int = 1024, seed: int = 1, extent: float = 1.0, size: float = 0.1):
    g = rng(seed)
    c = g.uniform(-extent, extent, size=(n, 1, 3)).astype(F32)
    off = g.uniform(-size, size, size=(n, 3, 3)).astype(F32)
    return (c + off).astype(F32).reshape(n, 9)


def torus(R: float = 1.0, r: float = 0.4, nu: int = 128, nv: int = 96, tilt_deg: float = 30.0):
    u = np.arange(nu + 1, dtype=np.float64) / nu * 2 * np.pi
    v = np.arange(nv + 1, dtype=np.float64) / nv * 2 * np.pi
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = (R + r * np.cos(vv)) * np.cos(uu)
    y = (R + r * np.cos(vv)) * np.sin(uu)
    z = r * np.sin(vv)
    a = np.deg2rad(tilt_deg)
    # tilt the ring about x so the camera on -z sees its face at an angle
    y2 = y * np.cos(a) - z * np.sin(a)
    z2 = y * np.sin(a) + z * np.cos(a)
    P = np.stack([x, y2, z2], axis=-1)
    p00 = P[:-1, :-1].reshape(-1, 3)
    p10 = P[1:, :-1].reshape(-1, 3)
    p11 = P[1:, 1:].reshape(-1, 3)
    p01 = P[:-1, 1:].reshape(-1, 3)
    t1 = np.concatenate([p00, p10, p11], axis=1)
    t2 = np.concatenate([p00, p11, p01], axis=1)
    return np.stack([t1, t2], axis=1).reshape(-1, 9).astype(F32)


def ground_plane(y: float = -0.6, half: float = 10.0):
    a = [-half, y, -half]
    b = [half, y, -half]
    c = [half, y, half]
    d = [-half, y, half]
    return np.array([a + b + c, a + c + d], dtype=F32)


def torus_scene(seed: int = 2):
    """~24.8k triangles: torus + ground plane + 200 small random triangles (edge <= 0.05)."""
    tor = torus()
    g = rng(seed)
    lo, hi = tor.reshape(-1, 3).min(0), tor.reshape(-1, 3).max(0)
    c = g.uniform(lo, hi, size=(200, 1, 3))
    off = g.uniform(-0.0125, 0.0125, size=(200, 3, 3))  # |edge| <= 0.025*sqrt(3) < 0.05
    small = (c + off).reshape(200, 9).astype(F32)
    return np.concatenate([tor, ground_plane(), small], axis=0).astype(F32)


def box_room(sx: float = 10.0, sy: float = 6.0, sz: float = 20.0):
    x0, x1, y0, y1, z0, z1 = -sx / 2, sx / 2, -sy / 2, sy / 2, -sz / 2, sz / 2
    P = lambda x, y, z: [x, y, z]
    quads = [
        (P(x0, y0, z0), P(x1, y0, z0), P(x1, y0, z1), P(x0, y0, z1)),  # floor
        (P(x0, y1, z0), P(x0, y1, z1), P(x1, y1, z1), P(x1, y1, z0)),  # ceiling
        (P(x0, y0, z0), P(x0, y0, z1), P(x0, y1, z1), P(x0, y1, z0)),  # left
        (P(x1, y0, z0), P(x1, y1, z0), P(x1, y1, z1), P(x1, y0, z1)),  # right
        (P(x0, y0, z0), P(x0, y1, z0), P(x1, y1, z0), P(x1, y0, z0)),  # back
        (P(x0, y0, z1), P(x1, y0, z1), P(x1, y1, z1), P(x0, y1, z1)),  # front
    ]
    tris = []
    for a, b, c, d in quads:
        tris.append(a + b + c)
        tris.append(a + c + d)
    return np.array(tris, dtype=F32)


def gaussian_clusters(nclusters: int, per: int, lo, hi, sigma: float, seed: int):
    g = rng(seed)
    c = g.uniform(lo, hi, size=(nclusters, 1, 1, 3))
    off = g.normal(0.0, sigma, size=(nclusters, per, 3, 3))
    return (c + off).reshape(-1, 9).astype(F32)


def room_scene(seed: int = 3, nclusters: int = 400, per: int = 200):
    room = box_room()
    cl = gaussian_clusters(nclusters, per, [-5, -3, -10], [5, 3, 10], 0.15, seed)
    return np.concatenate([room, cl], axis=0).astype(F32)


def cluster_scene(ntri: int, seed: int, per: int = 200):
    ncl = max(1, (ntri + per - 1) // per)
    v = gaussian_clusters(ncl, per, [-5, -5, -5], [5, 5, 5], 0.15, seed)
    return v[:ntri].copy()


# ---------------------------------------------------------------------------------------
# rays (dict of float32 arrays + width/height)

def pinhole_rays(W: int, H: int, origin=(0.0, 0.0, -3.0), fov_deg: float = 60.0, spp: int = 1,
                 seed: int | None = None):
    """Camera rays through pixel centres (or stratified jitter with spp > 1), samples of one
    pixel adjacent in x as the reference's sampler orders them: ray image (W*spp) x H."""
    tan = np.float32(np.tan(np.deg2rad(fov_deg) / 2))
    aspect = np.float32(W / H)
    ys, xs = np.meshgrid(np.arange(H, dtype=F32), np.arange(W, dtype=F32), indexing="ij")
    if spp == 1:
        jx = np.full((H, W, 1), 0.5, dtype=F32)
        jy = np.full((H, W, 1), 0.5, dtype=F32)
    else:
        g = rng(seed if seed is not None else 0)
        jx = g.random((H, W, spp), dtype=np.float32)
        jy = g.random((H, W, spp), dtype=np.float32)
    px = (((xs[..., None] + jx) / F32(W)) * F32(2) - F32(1)) * tan * aspect
    py = (F32(1) - ((ys[..., None] + jy) / F32(H)) * F32(2)) * tan
    pz = np.ones_like(px)
    l = np.sqrt((px * px + py * py) + pz * pz)
    inv = F32(1) / l
    dx, dy, dz = (px * inv).astype(F32), (py * inv).astype(F32), (pz * inv).astype(F32)
    n = W * H * spp
    o = np.asarray(origin, dtype=F32)
    return dict(ox=np.full(n, o[0], F32), oy=np.full(n, o[1], F32), oz=np.full(n, o[2], F32),
                dx=dx.reshape(-1), dy=dy.reshape(-1), dz=dz.reshape(-1),
                tmin=np.zeros(n, F32), tmax=np.full(n, INF, F32), valid=None,
                width=W * spp, height=H)


def random_rays(n: int, seed: int, lo=-5.0, hi=5.0, maxt: float = 10.0, width: int | None = None):
    """Random origins in a box, directions uniform on the sphere, tmax = maxt."""
    g = rng(seed)
    o = g.uniform(lo, hi, size=(n, 3)).astype(F32)
    d = g.normal(size=(n, 3)).astype(F32)
    l = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    d = (d * (F32(1) / l)[:, None]).astype(F32)
    W = width or n
    return dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=d[:, 0].copy(),
                dy=d[:, 1].copy(), dz=d[:, 2].copy(), tmin=np.zeros(n, F32),
                tmax=np.full(n, maxt, F32), valid=np.ones(n, np.uint8), width=W, height=n // W)


def shadow_rays_np(prim: dict, hit_id: np.ndarray, hit_t: np.ndarray, light):
    """numpy twin of rt_make_shadow_rays (pbrt SetSegment / PointLight::Sample_L)."""
    L = np.asarray(light, dtype=F32)
    ok = hit_id >= 0
    t = np.where(ok, hit_t, F32(0)).astype(F32)
    px = prim["ox"] + prim["dx"] * t
    py = prim["oy"] + prim["dy"] * t
    pz = prim["oz"] + prim["dz"] * t
    vx, vy, vz = L[0] - px, L[1] - py, L[2] - pz
    dist = np.sqrt((vx * vx + vy * vy) + vz * vz).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = (F32(1) / dist).astype(F32)
    z = lambda a: np.where(ok, a, F32(0)).astype(F32)
    return dict(ox=z(px), oy=z(py), oz=z(pz), dx=z(vx * inv), dy=z(vy * inv), dz=z(vz * inv),
                tmin=z(F32(1e-3) * t), tmax=z(dist * (F32(1) - F32(0))),
                valid=ok.astype(np.uint8), width=prim["width"], height=prim["height"])


def subimage(r: dict, x0: int, y0: int, w: int, h: int):
    """Crops a ray image to a w x h window (keeps the row-major image layout)."""
    W = r["width"]
    out = {}
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax", "valid"):
        a = r.get(k)
        out[k] = None if a is None else a.reshape(-1, W)[y0:y0 + h, x0:x0 + w].reshape(-1).copy()
    out["width"], out["height"] = w, h
    return out


def uneven_batch(nscenes: int = 64, seed0: int = 100):
    """The uneven batch: `nscenes` independent scenes.  Scene i (seed seed0 + i): triangle
    count log-uniform in [1e3, 3e5], Gaussian clusters of 200 triangles (sigma 0.15) with
    centres in [-5,5]^3; ray count log-uniform in [1.6e4, 1e6], half camera primaries (a
    square image, camera at (0,0,-12) looking +z, 50 deg FOV) and half random shadow rays
    (origins in [-5,5]^3, uniform directions, maxt 10) laid out 256 rays wide.  Image sides
    are rounded to multiples of 16 so that the default 4x4-leaf, height-3 hierarchy is exactly
    the reference's shape."""
    out = []
    for i in range(nscenes):
        g = rng(seed0 + i)
        ntri = int(round(np.exp(g.uniform(np.log(1e3), np.log(3e5)))))
        nray = int(round(np.exp(g.uniform(np.log(1.6e4), np.log(1e6)))))
        side = max(16, int(round(np.sqrt(nray / 2) / 16)) * 16)
        nsh = max(4096, int(round((nray - side * side) / 4096)) * 4096)
        out.append(dict(v9=cluster_scene(ntri, seed=seed0 + i),
                        primary=pinhole_rays(side, side, origin=(0.0, 0.0, -12.0), fov_deg=50.0),
                        shadow=random_rays(nsh, seed=10000 + seed0 + i, lo=-5.0, hi=5.0,
                                           maxt=10.0, width=256)))
    return out


WORKLOADS = {
    "random1k_naive_128": dict(scene="random1k", W=128, H=128, accel="naive", light=(2, 2, -2)),
    "torus_rayhierarchy_512": dict(scene="torus", W=512, H=512, accel="rayhierarchy",
                                   light=(2, 2, -2)),
    # one task of the 1024x1024 x 4 spp render: 256x256 pixels x 4 spp (1024x256 ray image)
    "room_raybvh_task": dict(scene="room", W=256, H=256, spp=4, seed=3, accel="ray-bvh",
                             origin=(0.0, 0.0, -9.0), fov=50.0, light=(1.0, 2.5, -8.0)),
    # the same task through the path integrator to depth 3: primary, 3 cosine-weighted
    # bounces, a shadow ray per hit at every depth
    "room_raybvh_path3": dict(scene="room", W=256, H=256, spp=4, seed=3, accel="ray-bvh",
                              origin=(0.0, 0.0, -9.0), fov=50.0, light=(1.0, 2.5, -8.0),
                              depth=3),
    # the whole 1024x1024 x 4 spp frame as the renderer's 16 tasks (4 x 4 tiles of 256x256
    # pixels, each a 1024x256 ray image), direct lighting / path depth 3
    "room_raybvh_frame": dict(scene="room", W=1024, H=1024, spp=4, seed=3, accel="ray-bvh",
                              origin=(0.0, 0.0, -9.0), fov=50.0, light=(1.0, 2.5, -8.0),
                              tiles=(4, 4)),
    "room_raybvh_frame_path3": dict(scene="room", W=1024, H=1024, spp=4, seed=3,
                                    accel="ray-bvh", origin=(0.0, 0.0, -9.0), fov=50.0,
                                    light=(1.0, 2.5, -8.0), tiles=(4, 4), depth=3),
}


def make_scene(name: str):
    if name == "random1k":
        return random_triangles(1024, seed=1)
    if name == "torus":
        return torus_scene(seed=2)
    if name == "room":
        return room_scene(seed=3)
    raise KeyError(name)
