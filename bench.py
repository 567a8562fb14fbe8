"""Benchmark: Mrays/s of the plain per-scene rayhierarchy accelerator on one MI355X.

Workload (default "torus_rayhierarchy_512"): the torus scene (24,778 triangles: UV torus
R=1 r=0.4 128x96 quads, ground plane, 200 small random triangles), a 512x512 pinhole camera
with one primary ray per pixel, 4x4-ray leaves, height 3, box_dir nodes, and one shadow ray
per hit toward a point light (directlighting).  One step = primary closest-hit trace (with
the fused hierarchy build and repair pass) + shadow-ray generation + shadow occlusion trace,
all on the device with inputs already resident in HBM.  Mrays/s counts primary rays plus
valid shadow rays.

Also reported: the dominant traversal kernel's HIP-event time against the HBM roofline
(algorithmic bytes of the reference's candidate set), and pbrt's CPU BVH (BVHAccel,
maxnodeprims 4, inclusive Triangle test) on the same rays on the host cores.
"""
import argparse
import importlib.util
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # MI355X dense FP32 (with packed FP32 instructions)
FP32_UNPACKED_TFLOPS = 78.6   # without them; this build compiles packed FP32 off


def _load(name, path, search=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[str(search)] if search else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters (FETCH_SIZE,
    WRITE_SIZE, calibrated against a plain copy) and recorded in profiles/*.md as a line
    `traffic-per-launch <workload> <kernel> <bytes> ...`; None when not profiled."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) >= 4 and parts[0] == "traffic-per-launch" and parts[1] == workload:
                pairs = dict(zip(parts[2::2], parts[3::2]))
                if kernel in pairs:
                    return int(float(pairs[kernel]))
    return None


def bench_uneven(args, rt, scenes, torch):
    """The uneven batch (64 scenes of log-uniform size, camera + random shadow rays): the
    plain per-scene form (every scene's closest and occlusion call in turn on one stream)
    against the batched form (one batch of 128 jobs, graph-replayed), same hits required."""
    print("[bench] uneven batch: generating 64 scenes", file=sys.stderr, flush=True)
    jobs = scenes.uneven_batch(64, seed0=100)
    ctx = rt.Context(0)
    stream = torch.cuda.Stream()
    scs, prims, shads = [], [], []
    for j in jobs:
        scs.append(rt.Scene(ctx, j["v9"], desc="rayhierarchy"))
        prims.append(rt.DeviceRays.from_numpy(j["primary"]))
        shads.append(rt.DeviceRays.from_numpy(j["shadow"]))
    n_prim = sum(p.count for p in prims)
    n_shad = sum(int(j["shadow"]["valid"].sum()) for j in jobs)
    rays_per_step = n_prim + n_shad
    hits_p = [rt.DeviceHits.empty(p.count) for p in prims]
    occ_p = [torch.empty(s.count, dtype=torch.uint8, device="cuda") for s in shads]

    def plain_step():
        for k in range(len(jobs)):
            scs[k].trace_closest(prims[k], rt.RT_CLOSEST_PRIMARY, hits=hits_p[k], stream=stream)
            scs[k].trace_occluded(shads[k], out=occ_p[k], stream=stream)

    b = rt.Batch(ctx)
    hits_b = [b.add_closest(scs[k], prims[k]) for k in range(len(jobs))]
    occ_b = [b.add_occluded(scs[k], shads[k]) for k in range(len(jobs))]
    if args.autotune:
        print("[bench] autotuning the batch", file=sys.stderr, flush=True)
        b.autotune(reps=1, stream=stream)
    b.set_param(rt.RT_PARAM_GRAPH, 1)

    def batch_step():
        b.run(stream=stream)

    def timed(fn, label):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        print(f"[bench] timing {args.steps} {label} steps", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        stream.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_plain = timed(plain_step, "plain")
    t_batch = timed(batch_step, "batched")
    # per-mode unit-kernel time of one eager batch run
    b.set_param(rt.RT_PARAM_GRAPH, 0)
    b.set_timing(True)
    b.run(stream=stream)
    stream.synchronize()
    k_closest = b.kernel_time(rt.RT_CLOSEST_PRIMARY)[0]
    k_occl = b.kernel_time(2)[0]
    split = {"closest_unit": round(b.kernel_time(rt.RT_CLOSEST_PRIMARY + rt.RT_TIME_UNIT)[0], 3),
             "closest_per_ray": round(b.kernel_time(rt.RT_CLOSEST_PRIMARY + rt.RT_TIME_PER_RAY)[0], 3),
             "occlusion_unit": round(b.kernel_time(2 + rt.RT_TIME_UNIT)[0], 3),
             "occlusion_per_ray": round(b.kernel_time(2 + rt.RT_TIME_PER_RAY)[0], 3)}
    sweep = {}
    for thr in args.ray_sweep:
        b.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, thr)
        b.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, thr)
        b.run(stream=stream)
        b.run(stream=stream)
        stream.synchronize()
        b.set_timing(True)
        b.run(stream=stream)
        stream.synchronize()
        sweep[thr] = {k: round(b.kernel_time(m)[0], 3) for k, m in (
            ("closest", 0), ("closest_per_ray", 0 + rt.RT_TIME_PER_RAY),
            ("occlusion", 2), ("occlusion_per_ray", 2 + rt.RT_TIME_PER_RAY))}
        b.set_timing(False)
    b.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, 0)
    b.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, 0)
    b.set_timing(False)
    mism_id = sum(int((hits_b[k].id != hits_p[k].id).sum().item()) for k in range(len(jobs)))
    mism_t = sum(int((hits_b[k].t.view(torch.int32) != hits_p[k].t.view(torch.int32)).sum().item())
                 for k in range(len(jobs)))
    mism_occ = sum(int((occ_b[k] != occ_p[k]).sum().item()) for k in range(len(jobs)))
    v_plain = rays_per_step / t_plain / 1e6
    v_batch = rays_per_step / t_batch / 1e6
    tuned = {name: b.get_param(p) for name, p in (
        ("clusters_closest", rt.RT_PARAM_CLUSTERS_CLOSEST),
        ("clusters_occluded", rt.RT_PARAM_CLUSTERS_OCCLUDED),
        ("grid_closest", rt.RT_PARAM_GRID_CLOSEST), ("grid_occluded", rt.RT_PARAM_GRID_OCCLUDED))}
    result = {
        "metric": "Mrays/s batched vs plain per-scene launch, same hits; p50/p99 request latency",
        "value": round(v_batch, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven_batch_64", "form": "batched (128 jobs, HIP graph)",
                   "scenes": len(jobs), "triangles": int(sum(j["v9"].shape[0] for j in jobs)),
                   "primary_rays": n_prim, "shadow_rays": n_shad, "accelerator": "rayhierarchy",
                   "leaf": "4x4", "height": 3, "node": "box_dir", "launch_shape": tuned},
        "plain": {"value": round(v_plain, 3), "ms_per_step": round(t_plain * 1e3, 3)},
        "batched_over_plain": round(v_batch / v_plain, 3),
        "kernels_ms": {"closest_traversal": round(k_closest, 3),
                       "occlusion_traversal": round(k_occl, 3), "split": split,
                       "ray_threshold_sweep": sweep},
        "parity": {"id_mismatches": mism_id, "t_mismatches": mism_t,
                   "occlusion_mismatches": mism_occ},
        "roofline": None, "cpu_baseline": None,
    }
    b.close()
    for sc in scs:
        sc.close()
    ctx.close()
    print(json.dumps(result))


def bench_frame(args, rt, scenes, torch, wl):
    """A whole frame as the renderer's tasks: the image is cut into tiles (one ray image of
    (tile width * spp) x tile height each).  Every pass (camera, then per depth the shadow and
    bounce rays) runs for all tasks either as one batch of jobs (batched form, HIP graph) or
    task by task as breadth-first.cpp runs them (plain form); both must give the same bits."""
    v9 = scenes.make_scene(wl["scene"])
    W, H, spp = wl["W"], wl["H"], wl.get("spp", 1)
    tx, ty = wl["tiles"]
    depth = wl.get("depth", 0)
    light = wl["light"]
    print(f"[bench] frame {W}x{H}x{spp}spp in {tx * ty} tasks, depth {depth}", file=sys.stderr,
          flush=True)
    full = scenes.pinhole_rays(W, H, origin=wl["origin"], fov_deg=wl["fov"], spp=spp,
                               seed=wl.get("seed"))
    tw, th = W * spp // tx, H // ty
    tasks = [scenes.subimage(full, x * tw, y * th, tw, th) for y in range(ty) for x in range(tx)]
    del full
    ctx = rt.Context(0)
    sc = rt.Scene(ctx, v9, desc=wl["accel"])
    stream = torch.cuda.Stream()
    T = len(tasks)
    prim = [rt.DeviceRays.from_numpy(r) for r in tasks]
    n = sum(p.count for p in prim)

    def buffers():
        """Device buffers of one form: hits, shadow rays and flags, bounce rays per depth."""
        hp = [rt.DeviceHits.empty(p.count) for p in prim]
        sh = [[None] * T for _ in range(depth + 1)]
        oc = [[torch.empty(p.count, dtype=torch.uint8, device="cuda") for p in prim]
              for _ in range(depth + 1)]
        bo = [[None] * T for _ in range(depth)]
        bh = [[rt.DeviceHits.empty(p.count) for p in prim] for _ in range(depth)]
        # shapes for every pass: run the chain once eagerly
        for k in range(T):
            sc.trace_closest(prim[k], rt.RT_CLOSEST_PRIMARY, hits=hp[k], stream=stream)
            sh[0][k] = rt.make_shadow_rays(prim[k], hp[k], light, stream=stream)
            src, srch = prim[k], hp[k]
            for d in range(depth):
                bo[d][k] = rt.make_bounce_rays(sc, src, srch, seed=17 + k, depth=d + 1,
                                               stream=stream)
                sc.trace_closest(bo[d][k], rt.RT_CLOSEST_SECONDARY, hits=bh[d][k], stream=stream)
                sh[d + 1][k] = rt.make_shadow_rays(bo[d][k], bh[d][k], light, stream=stream)
                src, srch = bo[d][k], bh[d][k]
        return hp, sh, oc, bo, bh

    P = buffers()
    B = buffers()

    def plain_step():
        hp, sh, oc, bo, bh = P
        for k in range(T):  # breadth-first.cpp: one task after the other
            sc.trace_closest(prim[k], rt.RT_CLOSEST_PRIMARY, hits=hp[k], stream=stream)
            rt.make_shadow_rays(prim[k], hp[k], light, stream=stream, out=sh[0][k])
            sc.trace_occluded(sh[0][k], out=oc[0][k], stream=stream)
            src, srch = prim[k], hp[k]
            for d in range(depth):
                rt.make_bounce_rays(sc, src, srch, seed=17 + k, depth=d + 1, stream=stream,
                                    out=bo[d][k])
                sc.trace_closest(bo[d][k], rt.RT_CLOSEST_SECONDARY, hits=bh[d][k], stream=stream)
                rt.make_shadow_rays(bo[d][k], bh[d][k], light, stream=stream, out=sh[d + 1][k])
                sc.trace_occluded(sh[d + 1][k], out=oc[d + 1][k], stream=stream)
                src, srch = bo[d][k], bh[d][k]

    # batched: one batch per pass, every task a job of it
    hp, sh, oc, bo, bh = B
    b_cam = rt.Batch(ctx)
    for k in range(T):
        b_cam.add_closest(sc, prim[k], rt.RT_CLOSEST_PRIMARY, hits=hp[k])
    b_occ = []
    for d in range(depth + 1):
        b = rt.Batch(ctx)
        for k in range(T):
            b.add_occluded(sc, sh[d][k], out=oc[d][k])
        b_occ.append(b)
    b_sec = []
    for d in range(depth):
        b = rt.Batch(ctx)
        for k in range(T):
            b.add_closest(sc, bo[d][k], rt.RT_CLOSEST_SECONDARY, hits=bh[d][k])
        b_sec.append(b)
    for b in [b_cam] + b_occ + b_sec:
        b.set_param(rt.RT_PARAM_GRAPH, 1)

    def batch_step():
        b_cam.run(stream=stream)
        for k in range(T):
            rt.make_shadow_rays(prim[k], hp[k], light, stream=stream, out=sh[0][k])
        b_occ[0].run(stream=stream)
        for d in range(depth):
            for k in range(T):
                src, srch = (prim[k], hp[k]) if d == 0 else (bo[d - 1][k], bh[d - 1][k])
                rt.make_bounce_rays(sc, src, srch, seed=17 + k, depth=d + 1, stream=stream,
                                    out=bo[d][k])
            b_sec[d].run(stream=stream)
            for k in range(T):
                rt.make_shadow_rays(bo[d][k], bh[d][k], light, stream=stream, out=sh[d + 1][k])
            b_occ[d + 1].run(stream=stream)

    def timed(fn, label):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        print(f"[bench] timing {args.steps} {label} steps", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        stream.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_plain = timed(plain_step, "plain")
    t_batch = timed(batch_step, "batched")
    n_shadow = sum(int(x.valid.sum().item()) for row in B[1] for x in row)
    n_bounce = sum(int(x.valid.sum().item()) for row in B[3] for x in row)
    rays_per_step = n + n_shadow + n_bounce

    def neq(a, b):
        return int((a != b).sum().item())
    mism_id = sum(neq(B[0][k].id, P[0][k].id) for k in range(T)) + sum(
        neq(B[4][d][k].id, P[4][d][k].id) for d in range(depth) for k in range(T))
    mism_t = sum(neq(B[0][k].t.view(torch.int32), P[0][k].t.view(torch.int32)) for k in range(T))
    mism_occ = sum(neq(B[2][d][k], P[2][d][k]) for d in range(depth + 1) for k in range(T))
    result = {
        "metric": "Mrays/s batched vs plain per-scene launch, same hits; p50/p99 request latency",
        "value": round(rays_per_step / t_batch / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "form": f"batched ({T} task jobs per pass, HIP graph)",
                   "accelerator": wl["accel"], "triangles": int(len(v9)),
                   "image": f"{W}x{H}x{spp}spp", "tasks": T, "task_ray_image": f"{tw}x{th}",
                   "primary_rays": n, "shadow_rays": n_shadow, "bounce_rays": n_bounce,
                   "path_depth": depth, "leaf": "4x4", "height": 3, "node": "box_dir"},
        "plain": {"value": round(rays_per_step / t_plain / 1e6, 3),
                  "ms_per_step": round(t_plain * 1e3, 3)},
        "batched_over_plain": round(t_plain / t_batch, 3),
        "parity": {"id_mismatches": mism_id, "t_mismatches": mism_t,
                   "occlusion_mismatches": mism_occ},
        "roofline": None, "cpu_baseline": None,
    }
    for b in [b_cam] + b_occ + b_sec:
        b.close()
    sc.close()
    ctx.close()
    print(json.dumps(result))


def bench_serving(rt, scenes, torch, rates, nreq, seed0=200):
    """Serving form: 16 resident cluster scenes (triangle counts log-uniform in [1e3, 1e6]),
    Poisson arrivals at each rate, resolutions uniform in {128, 256, 512, 1024} squared,
    1-4 spp, primary + shadow rays.  Returns p50 / p99 latency and aggregate Mrays/s per
    rate."""
    ctx = rt.Context(0)
    print("[bench] serving: building 16 resident scenes", file=sys.stderr, flush=True)
    scs = []
    for i in range(16):
        g = np.random.Generator(np.random.PCG64(seed0 + i))
        ntri = int(round(np.exp(g.uniform(np.log(1e3), np.log(1e6)))))
        scs.append(rt.Scene(ctx, scenes.cluster_scene(ntri, seed=seed0 + i), desc="rayhierarchy"))
    srv = rt.Server(ctx, scs, max_active=256)
    points = []
    try:
        for rate in rates:
            g = np.random.Generator(np.random.PCG64(seed0 + int(rate)))
            gaps = g.exponential(1.0 / rate, size=nreq)
            reqs = [dict(scene=int(g.integers(0, 16)), width=int(s), height=int(s),
                         spp=int(g.integers(1, 5)), origin=(0.0, 0.0, -12.0), fov_deg=50.0,
                         light=(0.0, 8.0, -8.0), seed=k)
                    for k, s in enumerate(g.choice([128, 256, 512, 1024], size=nreq))]
            print(f"[bench] serving: {nreq} requests at {rate:g} req/s", file=sys.stderr,
                  flush=True)
            t_start = time.perf_counter()
            due = t_start
            tickets = []
            for q, gap in zip(reqs, gaps):
                due += gap
                while True:
                    left = due - time.perf_counter()
                    if left <= 0:
                        break
                    time.sleep(min(left, 0.002))
                tickets.append(srv.submit(**q, outputs=True))
            res = [srv.wait(t, timeout_s=300.0) for t in tickets]
            t_end = time.perf_counter()
            lat = np.array([r["latency_s"] for r in res]) * 1e3
            rays = sum(int(r["primary_rays"]) + int((r["id"] >= 0).sum()) for r in res)
            first = min(r["submit_s"] for r in res)
            last = max(r["done_s"] for r in res)
            points.append({"rate_req_s": rate, "requests": nreq,
                           "p50_ms": round(float(np.percentile(lat, 50)), 2),
                           "p99_ms": round(float(np.percentile(lat, 99)), 2),
                           "mrays_s": round(rays / (last - first) / 1e6, 1),
                           "wall_s": round(t_end - t_start, 2)})
        ticks, most = srv.counters()
    finally:
        srv.close()
        for sc in scs:
            sc.close()
        ctx.close()
    return {"workload": "serving_16_scenes", "points": points, "ticks": ticks,
            "largest_tick_jobs": most}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="torus_rayhierarchy_512")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serve-rates", default="10,20,50,100,200,500",
                    help="request rates (req/s) of the serving leg; empty to skip it")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--autotune", action="store_true",
                    help="uneven_batch_64: autotune the batch's launch shape before timing")
    ap.add_argument("--ray-closest", type=int, default=0,
                    help="per-ray threshold (passing clusters) of closest-hit traces; 0 = built-in")
    ap.add_argument("--ray-occluded", type=int, default=0,
                    help="per-ray threshold of occlusion traces; 0 = built-in")
    ap.add_argument("--ray-sweep", type=lambda v: [int(x) for x in v.split(",") if x],
                    default=[], help="uneven_batch_64: per-ray thresholds to time")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed steps, run one plain device copy of this many MB "
                         "(calibrates profiler byte counters against a kernel of known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding is described in DESIGN.md and not launched here.", file=sys.stderr)
        sys.exit(2)

    import torch

    rt = _load("pbrt_v2_amd", ROOT / "pbrt-v2_amd" / "__init__.py", ROOT / "pbrt-v2_amd")
    scenes = _load("pbrt_v2_amd.scenes", ROOT / "pbrt-v2_amd" / "scenes.py")
    torch.cuda.set_device(0)
    if args.workload == "uneven_batch_64":
        bench_uneven(args, rt, scenes, torch)
        return
    wl = scenes.WORKLOADS[args.workload]
    if "tiles" in wl:
        bench_frame(args, rt, scenes, torch, wl)
        return
    print(f"[bench] workload {args.workload}: building scene", file=sys.stderr, flush=True)
    v9 = scenes.make_scene(wl["scene"])
    rays_np = scenes.pinhole_rays(wl["W"], wl["H"], origin=wl.get("origin", (0.0, 0.0, -3.0)),
                                  fov_deg=wl.get("fov", 60.0), spp=wl.get("spp", 1),
                                  seed=wl.get("seed"))
    light = wl["light"]
    accel = wl["accel"]

    ctx = rt.Context(0)
    sc = rt.Scene(ctx, v9, desc=accel)
    if args.ray_closest:
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, args.ray_closest)
    if args.ray_occluded:
        sc.set_param(rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED, args.ray_occluded)
    prim = rt.DeviceRays.from_numpy(rays_np)
    n = prim.count
    hits = rt.DeviceHits.empty(n)
    shadow = rt.make_shadow_rays(prim, sc.trace_closest(prim, rt.RT_CLOSEST_PRIMARY, hits=hits),
                                 light)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    # path bounces (cosine-weighted, device-generated), each with its own shadow rays
    depth = wl.get("depth", 0)
    bounces, bhits, bshadows, boccs = [], [], [], []
    src, src_hits = prim, hits
    for dpt in range(1, depth + 1):
        b = rt.make_bounce_rays(sc, src, src_hits, seed=17, depth=dpt)
        bh = sc.trace_closest(b, rt.RT_CLOSEST_SECONDARY)
        bounces.append(b)
        bhits.append(bh)
        bshadows.append(rt.make_shadow_rays(b, bh, light))
        boccs.append(torch.empty(n, dtype=torch.uint8, device="cuda"))
        src, src_hits = b, bh

    def step():
        sc.trace_closest(prim, rt.RT_CLOSEST_PRIMARY, hits=hits, stream=stream)
        rt.make_shadow_rays(prim, hits, light, stream=stream, out=shadow)
        sc.trace_occluded(shadow, out=occ, stream=stream)
        s_rays, s_hits = prim, hits
        for k in range(depth):
            rt.make_bounce_rays(sc, s_rays, s_hits, seed=17, depth=k + 1, stream=stream,
                                out=bounces[k])
            sc.trace_closest(bounces[k], rt.RT_CLOSEST_SECONDARY, hits=bhits[k], stream=stream)
            rt.make_shadow_rays(bounces[k], bhits[k], light, stream=stream, out=bshadows[k])
            sc.trace_occluded(bshadows[k], out=boccs[k], stream=stream)
            s_rays, s_hits = bounces[k], bhits[k]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_shadow = int(shadow.valid.sum().item())
    n_bounce = sum(int(b.valid.sum().item()) for b in bounces)
    n_bshadow = sum(int(b.valid.sum().item()) for b in bshadows)
    rays_per_step = n + n_shadow + n_bounce + n_bshadow

    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    sc.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    closest_ms, nc = sc.kernel_time(rt.RT_CLOSEST_PRIMARY)
    second_ms, ns = sc.kernel_time(rt.RT_CLOSEST_SECONDARY)
    occl_ms, no = sc.kernel_time(2)
    sc.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = rays_per_step * args.steps / (t1 - t0) / 1e6
    gpu_ids = hits.id.cpu().numpy()
    gpu_occ = occ.cpu().numpy()
    if args.copy_mb > 0:
        src = torch.empty(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    result = {
        "metric": "Mrays/s batched vs plain per-scene launch, same hits; p50/p99 request latency",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "form": "plain per-scene", "accelerator": accel,
                   "triangles": int(len(v9)),
                   "image": f"{wl['W']}x{wl['H']}x{wl.get('spp', 1)}spp",
                   "primary_rays": n, "shadow_rays": n_shadow + n_bshadow,
                   "bounce_rays": n_bounce, "path_depth": depth, "leaf": "4x4", "height": 3,
                   "node": "box_dir"},
        "kernels_ms": {"closest_traversal": round(closest_ms / max(nc, 1), 4),
                       "occlusion_traversal": round(occl_ms / max(no, 1), 4)},
    }
    if ns:
        result["kernels_ms"]["secondary_traversal"] = round(second_ms / ns, 4)

    roofline = None
    cpu = None
    if not args.no_cpu and depth == 0:
        orc = _load("rt_oracle_py", ROOT / "oracle" / "oracle.py")
        print("[bench] oracle counts + CPU baseline", file=sys.stderr, flush=True)
        p = orc.params(accel)
        if accel == "ray-bvh":
            nodes, roots, order = sc.handoff
            trace = lambda r, m: orc.trace_raybvh(v9, nodes, roots, order, r, p, m)
        else:
            trace = lambda r, m: orc.trace(v9, r, p, m)
        ref = trace(rays_np, orc.CLOSEST_PRIMARY)
        sh_np = scenes.shadow_rays_np(rays_np, ref["id"], ref["t"], light)
        oref = trace(sh_np, orc.OCCLUDED)
        result["parity"] = {"id_mismatches": int((gpu_ids != ref["id"]).sum()),
                            "occlusion_mismatches": int((gpu_occ != oref["occ"]).sum())}
        # Algorithmic bytes per launch (reference candidate set at the reference defaults):
        # rays 32 B + hit 16 B (closest) or 1 B (occlusion), 36 B per (leaf, triangle) pair,
        # 24 B per (hierarchy node, triangle) cull evaluation.
        cs, os_ = ref["stats"], oref["stats"]
        # a ray-bvh cull evaluation reads one hierarchy node (24 B) and one GPUNode (32 B)
        per_cull = 56 if accel == "ray-bvh" else 24
        bytes_closest = n * 48 + cs.leaf_pairs * 36 + cs.cull_tests * per_cull
        bytes_occl = n_shadow * 33 + os_.leaf_pairs * 36 + os_.cull_tests * per_cull
        per_c = closest_ms / max(nc, 1)
        per_o = occl_ms / max(no, 1)
        if per_o >= per_c:
            name, b, t = "occlusion_traversal", bytes_occl, per_o
        else:
            name, b, t = "closest_traversal", bytes_closest, per_c
        achieved = b / (t * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": profiled_traffic(args.workload, name), "bytes_per_launch": int(b),
                    "counts": {"cull_tests": int((os_ if name[0] == "o" else cs).cull_tests),
                               "leaf_pairs": int((os_ if name[0] == "o" else cs).leaf_pairs)}}
        # Arithmetic side: ~45 FLOP per ray-triangle leaf test over the oracle's leaf tests
        # (valid rays of every reached (leaf, triangle) pair; occlusion stops testing a ray
        # once the oracle finds it occluded).  The cull evaluations are not counted.
        ops = 45 * int((os_ if name[0] == "o" else cs).ray_tests)
        tflops = ops / (t * 1e-3) / 1e12
        roofline["valu"] = {"flops_per_launch": ops, "achieved": round(tflops, 3),
                            "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
                            "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                            "peak_unpacked": FP32_UNPACKED_TFLOPS}
        # CPU baseline: pbrt's BVHAccel on the same primary + shadow rays, all host threads
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        bvh = orc.CpuBvh(v9, max_prims=4)
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            bvh.closest(rays_np, nthreads=threads)
            bvh.occluded(sh_np, nthreads=threads)
            times.append(time.perf_counter() - c0)
        # one core, for the per-core rate
        times1 = []
        for _ in range(3):
            c0 = time.perf_counter()
            bvh.closest(rays_np, nthreads=1)
            bvh.occluded(sh_np, nthreads=1)
            times1.append(time.perf_counter() - c0)
        bvh.close()
        med = statistics.median(times)
        nrays = n + int(sh_np["valid"].sum())
        cpu = {"value": round(nrays / med / 1e6, 3), "unit": "Mrays/s",
               "cores": threads, "kind": "port",
               "sample": f"full workload (primary + shadow rays), median of 5 runs, "
                         f"pbrt BVHAccel SAH maxnodeprims 4 restated in oracle/",
               "value_1core": round(nrays / statistics.median(times1) / 1e6, 3)}
    result["roofline"] = roofline
    result["cpu_baseline"] = cpu
    if args.serve_rates:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        result["serving"] = bench_serving(rt, scenes, torch, rates, args.serve_requests)
    sc.close()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
