"""Benchmark: batched vs plain per-scene Mrays/s of the MI355X ray-space accelerators, with
p50/p99 request latency of the serving form.

Default workload "uneven_batch_64": 64 independent scenes of uneven size (triangle counts
log-uniform in [1e3, 3e5], seeded Gaussian triangle clusters in [-5,5]^3), each with a camera
ray image and a set of random-direction shadow segments (ray counts log-uniform in [1.6e4,
1e6], half each), rayhierarchy accelerator at the reference defaults (4x4-ray leaves, height 3,
box_dir nodes, repairRun 1).  One step traces every scene's closest hits and occlusion:
  - batched: one batch of 128 jobs, launch shape picked by the autotuner, replayed as a HIP
    graph (the line's `value`);
  - plain: the 64 per-scene call pairs in turn, rays resident in HBM (`plain`), and the same
    with every call's rays copied host-to-device first (`plain_h2d`).
All three must give the same bits; sampled ray blocks are checked against the oracle, with
near ties, ulp-marginal rays and differences to the exact pbrt predicate counted.

Also reported: the roofline of the dominant traversal kernel (algorithmic bytes from the
oracle's restatement of its work, HIP-event time of the kernel), pbrt's CPU BVH on a bounded
sample of the same rays, the plain rayhierarchy torus workload, and the serving form at
Poisson request rates.
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
METRIC = "Mrays/s batched vs plain per-scene launch, same hits; p50/p99 request latency"


def _load(name, path, search=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[str(search)] if search else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profile_facts(workload, kernel):
    """Facts about `kernel` on `workload` recorded in profiles/*.md by the rocprofv3 runs, as
    lines `<fact> <workload> <kernel> <value>`:
      traffic-per-launch  HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by what the
                          same counters read for a plain copy of known size)
      valu-busy           fraction of the launch the VALU was busy (SQ counters)
      lane-utilisation    fraction of the VALU lanes active while it issued (SQ counters)
    Returns {fact: value} (empty when not profiled)."""
    out = {}
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) >= 4 and parts[1] == workload and parts[2] == kernel:
                try:
                    out[parts[0]] = float(parts[3])
                except ValueError:
                    pass
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def timed(fn, steps, warmup, stream, label):
    for _ in range(warmup):
        fn()
    stream.synchronize()
    log(f"timing {steps} {label} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    stream.synchronize()
    return (time.perf_counter() - t0) / steps


def bound_of(facts, hbm_meas, hbm_alg, valu_frac):
    """The roof that limits the kernel.  From the profiled counters when they are recorded:
    HBM when the measured DRAM traffic uses a larger share of its peak than the VALU is busy,
    else VALU.  Without them, the same comparison of the algorithmic model's shares."""
    busy = facts.get("valu-busy")
    if busy is not None and hbm_meas is not None:
        return "hbm" if hbm_meas >= busy else "valu"
    return "hbm" if hbm_alg >= valu_frac else "valu"


def per_ray_roofline(name, t_ms, st, mode, workload):
    """Roofline of a per-ray kernel launch from the oracle's restatement of its work: per
    traced ray 32 B (o, d, tmin, tmax) + 1 B valid + the result (16 B id/t/u/v, 1 B occlusion),
    64 B per search-BVH node loaded, 52 B per strict leaf test (v1, e1, e2 and the triangle
    id), 64 B per chain cull (one hierarchy node) and 36 B per chain (the triangle's
    vertices).  Arithmetic: 24 FLOP per node (two slab tests), 45 per leaf test, 60 per cull."""
    out_b = 1 if mode == 2 else 16
    b = (st.rays * (33 + out_b) + st.visits * 64 + st.tests * 52 + st.culls * 64
         + st.chains * 36)
    achieved = b / (t_ms * 1e-3) / 1e9
    flops = 24 * st.visits + 45 * st.tests + 60 * st.culls
    tflops = flops / (t_ms * 1e-3) / 1e12
    facts = profile_facts(workload, name)
    traffic = facts.get("traffic-per-launch")
    hbm_meas = None if traffic is None else traffic / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    return {"bound": bound_of(facts, hbm_meas, achieved / HBM_PEAK_GBS,
                              tflops / FP32_UNPACKED_TFLOPS), "kernel": name,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else int(traffic),
            "bytes_per_launch": int(b), "ms_per_launch": round(t_ms, 4),
            "counts": {"rays": int(st.rays), "node_loads": int(st.visits),
                       "leaf_tests": int(st.tests), "chain_culls": int(st.culls),
                       "chains": int(st.chains)},
            "valu": {"flops_per_launch": int(flops), "achieved": round(tflops, 3),
                     "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                     "peak_unpacked": FP32_UNPACKED_TFLOPS,
                     "busy_measured": facts.get("valu-busy"),
                     "lane_utilisation_measured": facts.get("lane-utilisation")}}


def _window(r, x0, y0, w, h):
    out = {}
    W = r["width"]
    for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax", "valid"):
        a = r.get(k)
        out[k] = None if a is None else a.reshape(-1, W)[y0:y0 + h, x0:x0 + w].reshape(-1).copy()
    out["width"], out["height"] = w, h
    return out


def bench_uneven(args, rt, scenes, torch, orc):
    """The uneven batch: batched vs plain per-scene launches, same hits required."""
    log("uneven batch: generating 64 scenes")
    jobs = scenes.uneven_batch(64, seed0=100)
    ctx = rt.Context(0)
    stream = torch.cuda.Stream()
    scs, prims, shads = [], [], []
    for j in jobs:
        scs.append(rt.Scene(ctx, j["v9"], desc="rayhierarchy"))
        prims.append(rt.DeviceRays.from_numpy(j["primary"]))
        shads.append(rt.DeviceRays.from_numpy(j["shadow"]))
    S = len(jobs)
    n_prim = sum(p.count for p in prims)
    n_shad = sum(int(j["shadow"]["valid"].sum()) for j in jobs)
    rays_per_step = n_prim + n_shad
    hits_p = [rt.DeviceHits.empty(p.count) for p in prims]
    occ_p = [torch.empty(s.count, dtype=torch.uint8, device="cuda") for s in shads]

    def plain_step():
        for k in range(S):
            scs[k].trace_closest(prims[k], rt.RT_CLOSEST_PRIMARY, hits=hits_p[k], stream=stream)
            scs[k].trace_occluded(shads[k], out=occ_p[k], stream=stream)

    # the plain form as the reference's Intersect / IntersectP calls take it: rays come from
    # host memory with every call (pinned staging copies of the same rays)
    fields = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax", "valid")

    def pinned(d):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k])).pin_memory() for k in fields
                if d.get(k) is not None}
    h_prim = [pinned(j["primary"]) for j in jobs]
    h_shad = [pinned(j["shadow"]) for j in jobs]

    def plain_h2d_step():
        with torch.cuda.stream(stream):
            for k in range(S):
                for f, t in h_prim[k].items():
                    getattr(prims[k], f).copy_(t, non_blocking=True)
                scs[k].trace_closest(prims[k], rt.RT_CLOSEST_PRIMARY, hits=hits_p[k],
                                     stream=stream)
                for f, t in h_shad[k].items():
                    getattr(shads[k], f).copy_(t, non_blocking=True)
                scs[k].trace_occluded(shads[k], out=occ_p[k], stream=stream)

    b = rt.Batch(ctx)
    hits_b = [b.add_closest(scs[k], prims[k]) for k in range(S)]
    occ_b = [b.add_occluded(scs[k], shads[k]) for k in range(S)]
    if not args.no_autotune:
        log("autotuning the batch")
        b.autotune(reps=2, stream=stream)
    b.set_param(rt.RT_PARAM_GRAPH, 1)

    def batch_step():
        b.run(stream=stream)

    t_plain = timed(plain_step, args.steps, args.warmup, stream, "plain")
    t_h2d = timed(plain_h2d_step, args.steps, args.warmup, stream, "plain (rays from host)")
    t_batch = timed(batch_step, args.steps, args.warmup, stream, "batched")
    tuned = {name: b.get_param(p) for name, p in (
        ("clusters_closest", rt.RT_PARAM_CLUSTERS_CLOSEST),
        ("clusters_occluded", rt.RT_PARAM_CLUSTERS_OCCLUDED),
        ("grid_closest", rt.RT_PARAM_GRID_CLOSEST), ("grid_occluded", rt.RT_PARAM_GRID_OCCLUDED),
        ("ray_clusters_closest", rt.RT_PARAM_RAY_CLUSTERS_CLOSEST),
        ("ray_clusters_occluded", rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED),
        ("ray_grid_closest", rt.RT_PARAM_RAY_GRID_CLOSEST),
        ("ray_grid_occluded", rt.RT_PARAM_RAY_GRID_OCCLUDED),
        ("walk_closest", rt.RT_PARAM_WALK_CLOSEST), ("walk_occluded", rt.RT_PARAM_WALK_OCCLUDED),
        ("tile_closest", rt.RT_PARAM_TILE_CLOSEST), ("tile_occluded", rt.RT_PARAM_TILE_OCCLUDED))}

    # per-kernel time of one eager run (unit kernel and per-ray kernel of each mode)
    def eager_split():
        b.set_param(rt.RT_PARAM_GRAPH, 0)
        b.set_timing(True)
        b.run(stream=stream)
        stream.synchronize()
        sp = {k: b.kernel_time(m)[0] for k, m in (
            ("closest_unit", rt.RT_CLOSEST_PRIMARY + rt.RT_TIME_UNIT),
            ("closest_per_ray", rt.RT_CLOSEST_PRIMARY + rt.RT_TIME_PER_RAY),
            ("occlusion_unit", 2 + rt.RT_TIME_UNIT), ("occlusion_per_ray", 2 + rt.RT_TIME_PER_RAY),
            ("closest", rt.RT_CLOSEST_PRIMARY), ("occlusion", 2))}
        b.set_timing(False)
        return sp
    split = eager_split()

    # parity: batched vs plain, bit for bit
    def neq(a, c):
        return int((a != c).sum().item())
    mism_id = sum(neq(hits_b[k].id, hits_p[k].id) for k in range(S))
    mism_t = sum(neq(hits_b[k].t.view(torch.int32), hits_p[k].t.view(torch.int32))
                 for k in range(S))
    mism_occ = sum(neq(occ_b[k], occ_p[k]) for k in range(S))

    # parity against the oracle on sampled ray blocks (a 16x16 block is a whole hierarchy
    # subtree, so the oracle's trace of the block's window is the reference's answer for it)
    log("oracle on sampled blocks")
    g = np.random.default_rng(7)
    p = orc.params("rayhierarchy")
    smp = {"blocks": 0, "rays": 0, "id_mismatches": 0, "occlusion_mismatches": 0,
           "near_ties": 0, "ulp_marginal": 0, "oref_vs_oexact_id": 0,
           "oref_vs_oexact_occlusion": 0}
    for k in range(0, S, max(1, S // args.oracle_scenes)):
        v9 = jobs[k]["v9"]
        ex = orc.CpuBvh(v9, max_prims=4)
        for kind in ("primary", "shadow"):
            r = jobs[k][kind]
            bx = int(g.integers(0, r["width"] // 16)) * 16
            by = int(g.integers(0, r["height"] // 16)) * 16
            win = _window(r, bx, by, 16, 16)
            sel = (np.arange(by, by + 16)[:, None] * r["width"] + np.arange(bx, bx + 16)).ravel()
            smp["blocks"] += 1
            smp["rays"] += 256
            if kind == "primary":
                ref = orc.trace(v9, win, p, orc.CLOSEST_PRIMARY, flags=True)
                gid = hits_b[k].id.cpu().numpy()[sel]
                smp["id_mismatches"] += int((gid != ref["id"]).sum())
                smp["near_ties"] += int((ref["flags"] & orc.NEAR_TIE != 0).sum())
                smp["oref_vs_oexact_id"] += int((ex.closest(win)["id"] != ref["id"]).sum())
            else:
                ref = orc.trace(v9, win, p, orc.OCCLUDED, flags=True)
                gocc = occ_b[k].cpu().numpy()[sel]
                smp["occlusion_mismatches"] += int((gocc != ref["occ"]).sum())
                smp["oref_vs_oexact_occlusion"] += int((ex.occluded(win) != ref["occ"]).sum())
            smp["ulp_marginal"] += int((ref["flags"] & orc.ULP_MARGINAL != 0).sum())
        ex.close()

    # rooflines of both per-ray kernels: a kernel's work is every ray it traces, so each is timed
    # with every block of its mode sent to it (median of 3 eager runs), and the oracle's
    # restatement counts that work; the line's roofline is the one whose kernel takes longer
    # in the step, the other is reported beside it
    b.set_param(rt.RT_PARAM_GRAPH, 0)
    rooflines = {}
    for mname, hm in (("closest", rt.RT_CLOSEST_PRIMARY), ("occlusion", 2)):
        pr = rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED if hm == 2 else rt.RT_PARAM_RAY_CLUSTERS_CLOSEST
        old = b.get_param(pr)
        b.set_param(pr, 1)
        b.run(stream=stream)
        times = []
        for _ in range(3):
            b.set_timing(True)
            b.run(stream=stream)
            stream.synchronize()
            times.append(b.kernel_time(hm + rt.RT_TIME_PER_RAY)[0])
            b.set_timing(False)
        t_kernel = statistics.median(times)
        b.set_param(pr, old)
        log(f"oracle counts of the per-ray kernel's work ({mname})")
        tot = None
        for k in range(S):
            r = jobs[k]["shadow" if hm == 2 else "primary"]
            st = orc.search_trace(jobs[k]["v9"], r, orc.params("rayhierarchy"), hm,
                                  nthreads=host_threads())["stats"]
            if tot is None:
                tot = st
            else:
                for f in ("rays", "visits", "tests", "culls", "chains"):
                    setattr(tot, f, getattr(tot, f) + getattr(st, f))
        kname = f"ray_path_kernel<{hm},0,false>"
        rooflines[mname] = per_ray_roofline(kname, t_kernel, tot, hm, "uneven_batch_64")
        rooflines[mname]["measured_as"] = (
            f"the {mname} jobs of the batch with every block sent to the per-ray kernel, "
            f"median of 3 eager runs")
    heavy = "occlusion" if split["occlusion_per_ray"] >= split["closest_per_ray"] else "closest"
    roofline = rooflines[heavy]
    roofline_other = rooflines["closest" if heavy == "occlusion" else "occlusion"]

    # CPU baseline: pbrt's BVHAccel (SAH, maxnodeprims 4, inclusive Triangle test) on the
    # camera and shadow rays of every second scene, all host threads, median of 3; one core on
    # every eighth scene
    log("CPU baseline")
    cpu = None
    if not args.no_cpu:
        thr = host_threads()
        sample = list(range(0, S, 2))
        bvhs = [orc.CpuBvh(jobs[k]["v9"], max_prims=4) for k in sample]
        nr = sum(jobs[k]["primary"]["width"] * jobs[k]["primary"]["height"]
                 + int(jobs[k]["shadow"]["valid"].sum()) for k in sample)
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            for bv, k in zip(bvhs, sample):
                bv.closest(jobs[k]["primary"], nthreads=thr)
                bv.occluded(jobs[k]["shadow"], nthreads=thr)
            ts.append(time.perf_counter() - c0)
        one = list(range(0, len(sample), 4))
        nr1 = sum(jobs[sample[i]]["primary"]["width"] * jobs[sample[i]]["primary"]["height"]
                  + int(jobs[sample[i]]["shadow"]["valid"].sum()) for i in one)
        c0 = time.perf_counter()
        for i in one:
            bvhs[i].closest(jobs[sample[i]]["primary"], nthreads=1)
            bvhs[i].occluded(jobs[sample[i]]["shadow"], nthreads=1)
        t1 = time.perf_counter() - c0
        for bv in bvhs:
            bv.close()
        cpu = {"value": round(nr / statistics.median(ts) / 1e6, 3), "unit": "Mrays/s",
               "cores": thr, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"{len(sample)} of the 64 scenes (every second), their {nr} camera + "
                         f"shadow rays, median of 3 runs; BVH build excluded",
               "value_1core": round(nr1 / t1 / 1e6, 3),
               "sample_1core": f"{len(one)} of those scenes, {nr1} rays, one run"}

    result = {
        "metric": METRIC, "value": round(rays_per_step / t_batch / 1e6, 3), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven_batch_64",
                   "form": f"batched ({2 * S} jobs, HIP graph with the closest and occlusion "
                           "groups as parallel branches, autotuned launch shape)",
                   "scenes": S, "triangles": int(sum(j["v9"].shape[0] for j in jobs)),
                   "primary_rays": n_prim, "shadow_rays": n_shad, "accelerator": "rayhierarchy",
                   "leaf": "4x4", "height": 3, "node": "box_dir", "launch_shape": tuned},
        "plain": {"value": round(rays_per_step / t_plain / 1e6, 3),
                  "ms_per_step": round(t_plain * 1e3, 3), "rays": "resident in HBM"},
        "plain_h2d": {"value": round(rays_per_step / t_h2d / 1e6, 3),
                      "ms_per_step": round(t_h2d * 1e3, 3),
                      "rays": "copied host-to-device (pinned) with every call"},
        "batched_over_plain": round(t_plain / t_batch, 3),
        "batched_over_plain_h2d": round(t_h2d / t_batch, 3),
        "kernels_ms": {k: round(v, 3) for k, v in split.items()},
        "parity": {"batched_vs_plain": {"id_mismatches": mism_id, "t_mismatches": mism_t,
                                        "occlusion_mismatches": mism_occ},
                   "oracle_sample": smp},
        "roofline": roofline, "roofline_other_per_ray_kernel": roofline_other,
        "cpu_baseline": cpu,
    }
    b.close()
    for sc in scs:
        sc.close()
    ctx.close()
    return result


def bench_frame(args, rt, scenes, torch, orc, wl):
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

    # batched: every task a job of one batch per pass. A depth's shadow rays and the next
    # depth's bounce rays both start from that depth's hits, so they share a batch: its
    # occlusion and closest-hit groups run as parallel branches of the batch's graph.
    hp, sh, oc, bo, bh = B
    b_cam = rt.Batch(ctx)
    for k in range(T):
        b_cam.add_closest(sc, prim[k], rt.RT_CLOSEST_PRIMARY, hits=hp[k])
    b_pass = []
    for d in range(depth + 1):
        b = rt.Batch(ctx)
        for k in range(T):
            b.add_occluded(sc, sh[d][k], out=oc[d][k])
        if d < depth:
            for k in range(T):
                b.add_closest(sc, bo[d][k], rt.RT_CLOSEST_SECONDARY, hits=bh[d][k])
        b_pass.append(b)
    for b in [b_cam] + b_pass:
        b.set_param(rt.RT_PARAM_GRAPH, 1)

    def batch_step():
        b_cam.run(stream=stream)
        for d in range(depth + 1):
            for k in range(T):
                src, srch = (prim[k], hp[k]) if d == 0 else (bo[d - 1][k], bh[d - 1][k])
                rt.make_shadow_rays(src, srch, light, stream=stream, out=sh[d][k])
                if d < depth:
                    rt.make_bounce_rays(sc, src, srch, seed=17 + k, depth=d + 1, stream=stream,
                                        out=bo[d][k])
            b_pass[d].run(stream=stream)

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

    # CPU baseline: pbrt's BVHAccel (SAH, maxnodeprims 4, inclusive Triangle test) on the first
    # task's rays of every pass (camera, bounce and shadow rays as the batched form made them),
    # all host threads, median of 3; BVH build excluded
    cpu = None
    if not args.no_cpu:
        def host(d):
            out = {k: getattr(d, k).cpu().numpy() for k in
                   ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")}
            out["valid"] = d.valid.cpu().numpy() if d.valid is not None else None
            out["width"], out["height"] = d.width, d.height
            return out
        closest_s = [tasks[0]] + [host(B[3][d][0]) for d in range(depth)]
        shadow_s = [host(B[1][d][0]) for d in range(depth + 1)]
        nr = sum(int(r["width"]) * int(r["height"]) if r.get("valid") is None
                 else int(np.asarray(r["valid"]).sum()) for r in closest_s) + \
            sum(int(r["valid"].sum()) for r in shadow_s)
        bv = orc.CpuBvh(v9, max_prims=4)
        thr = host_threads()
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            for r in closest_s:
                bv.closest(r, nthreads=thr)
            for r in shadow_s:
                bv.occluded(r, nthreads=thr)
            ts.append(time.perf_counter() - c0)
        bv.close()
        cpu = {"value": round(nr / statistics.median(ts) / 1e6, 3), "unit": "Mrays/s",
               "cores": thr, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"task 0 of {T}: its camera, bounce and shadow rays ({nr} rays), median "
                         f"of 3 runs; BVH build excluded"}
    result = {
        "metric": "Mrays/s batched vs plain per-scene launch, same hits; p50/p99 request latency",
        "value": round(rays_per_step / t_batch / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload,
                   "form": f"batched ({T} task jobs per pass, HIP graph; a depth's shadow "
                           "and bounce rays in one batch)",
                   "accelerator": wl["accel"], "triangles": int(len(v9)),
                   "image": f"{W}x{H}x{spp}spp", "tasks": T, "task_ray_image": f"{tw}x{th}",
                   "primary_rays": n, "shadow_rays": n_shadow, "bounce_rays": n_bounce,
                   "path_depth": depth, "leaf": "4x4", "height": 3, "node": "box_dir"},
        "plain": {"value": round(rays_per_step / t_plain / 1e6, 3),
                  "ms_per_step": round(t_plain * 1e3, 3)},
        "batched_over_plain": round(t_plain / t_batch, 3),
        "parity": {"id_mismatches": mism_id, "t_mismatches": mism_t,
                   "occlusion_mismatches": mism_occ},
        "roofline": None, "cpu_baseline": cpu,
    }
    for b in [b_cam] + b_pass:
        b.close()
    sc.close()
    ctx.close()
    print(json.dumps(result))



def bench_plain(args, rt, scenes, torch, orc, name):
    """One plain per-scene workload (a scene of WORKLOADS): the primary trace, shadow rays and
    occlusion trace (and path bounces to its depth) on the device; the dominant traversal
    kernel against the HBM roofline by the reference candidate set's algorithmic bytes."""
    wl = scenes.WORKLOADS[name]
    log(f"workload {name}: building scene")
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
    depth = wl.get("depth", 0)
    bounces, bhits, bshadows, boccs = [], [], [], []
    src, src_hits = prim, hits
    for dpt in range(1, depth + 1):
        bo = rt.make_bounce_rays(sc, src, src_hits, seed=17, depth=dpt)
        bh = sc.trace_closest(bo, rt.RT_CLOSEST_SECONDARY)
        bounces.append(bo)
        bhits.append(bh)
        bshadows.append(rt.make_shadow_rays(bo, bh, light))
        boccs.append(torch.empty(n, dtype=torch.uint8, device="cuda"))
        src, src_hits = bo, bh

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
    n_bounce = sum(int(x.valid.sum().item()) for x in bounces)
    n_bshadow = sum(int(x.valid.sum().item()) for x in bshadows)
    rays_per_step = n + n_shadow + n_bounce + n_bshadow
    log(f"timing {args.steps} steps")
    sc.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    km = {}
    for key, m in (("closest", rt.RT_CLOSEST_PRIMARY), ("secondary", rt.RT_CLOSEST_SECONDARY),
                   ("occlusion", 2)):
        ms, cnt = sc.kernel_time(m)
        if cnt:
            km[key] = ms / cnt
            if accel != "naive":
                km[key + "_unit"] = sc.kernel_time(m + rt.RT_TIME_UNIT)[0] / cnt
                km[key + "_per_ray"] = sc.kernel_time(m + rt.RT_TIME_PER_RAY)[0] / cnt
    sc.set_timing(False)
    gpu_ids = hits.id.cpu().numpy()
    gpu_occ = occ.cpu().numpy()
    result = {
        "metric": METRIC, "value": round(rays_per_step * args.steps / (t1 - t0) / 1e6, 3),
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": name, "form": "plain per-scene", "accelerator": accel,
                   "triangles": int(len(v9)),
                   "image": f"{wl['W']}x{wl['H']}x{wl.get('spp', 1)}spp",
                   "primary_rays": n, "shadow_rays": n_shadow + n_bshadow,
                   "bounce_rays": n_bounce, "path_depth": depth, "leaf": "4x4", "height": 3,
                   "node": "box_dir"},
        "kernels_ms": {k: round(v, 4) for k, v in km.items()},
    }
    roofline = cpu = None
    if depth == 0 and not args.no_cpu:
        log("oracle counts + CPU baseline")
        p = orc.params(accel)
        if accel == "ray-bvh":
            nodes, roots, order = sc.handoff
            trace = lambda r, m: orc.trace_raybvh(v9, nodes, roots, order, r, p, m, flags=True)
        else:
            trace = lambda r, m: orc.trace(v9, r, p, m, flags=True)
        ref = trace(rays_np, orc.CLOSEST_PRIMARY)
        sh_np = scenes.shadow_rays_np(rays_np, ref["id"], ref["t"], light)
        oref = trace(sh_np, orc.OCCLUDED)
        result["parity"] = {"id_mismatches": int((gpu_ids != ref["id"]).sum()),
                            "occlusion_mismatches": int((gpu_occ != oref["occ"]).sum()),
                            "near_ties": int((ref["flags"] & orc.NEAR_TIE != 0).sum()),
                            "ulp_marginal": int((ref["flags"] & orc.ULP_MARGINAL != 0).sum()
                                                + (oref["flags"] & orc.ULP_MARGINAL != 0).sum())}
        if accel != "naive":
            # the block path's work: rays 32 B + hit 16 B (closest) or 1 B (occlusion), 36 B per
            # (leaf, triangle) pair reaching the leaf test, 24 B per (hierarchy node, triangle)
            # cull evaluation (+ 32 B per BVH node for ray-bvh) of the reference candidate set
            cs, os_ = ref["stats"], oref["stats"]
            per_cull = 56 if accel == "ray-bvh" else 24
            bc = n * 48 + cs.leaf_pairs * 36 + cs.cull_tests * per_cull
            bo_ = n_shadow * 33 + os_.leaf_pairs * 36 + os_.cull_tests * per_cull
            tc, to = km.get("closest_unit", 0.0), km.get("occlusion_unit", 0.0)
            # the unit kernels of this single-job launch: 256-ray blocks, 3 levels
            name_k, b, t, st = (("rh_unit_kernel<2,0,256,3>", bo_, to, os_) if to >= tc
                                else ("rh_unit_kernel<0,0,256,3>", bc, tc, cs))
            if t > 0:
                achieved = b / (t * 1e-3) / 1e9
                ops = 45 * int(st.ray_tests)
                tfl = ops / (t * 1e-3) / 1e12
                facts = profile_facts(name, name_k)
                traffic = facts.get("traffic-per-launch")
                hbm_meas = None if traffic is None else traffic / (t * 1e-3) / 1e9 / HBM_PEAK_GBS
                roofline = {"bound": bound_of(facts, hbm_meas, achieved / HBM_PEAK_GBS,
                                              tfl / FP32_UNPACKED_TFLOPS),
                            "kernel": name_k, "achieved": round(achieved, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": None if traffic is None else int(traffic),
                            "bytes_per_launch": int(b),
                            "counts": {"cull_tests": int(st.cull_tests),
                                       "leaf_pairs": int(st.leaf_pairs)},
                            "valu": {"flops_per_launch": ops, "achieved": round(tfl, 3),
                                     "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
                                     "frac": round(tfl / FP32_PEAK_TFLOPS, 4),
                                     "peak_unpacked": FP32_UNPACKED_TFLOPS,
                                     "busy_measured": facts.get("valu-busy"),
                                     "lane_utilisation_measured":
                                         facts.get("lane-utilisation")}}
        thr = host_threads()
        bvh = orc.CpuBvh(v9, max_prims=4)
        ts = []
        for _ in range(5):
            c0 = time.perf_counter()
            bvh.closest(rays_np, nthreads=thr)
            bvh.occluded(sh_np, nthreads=thr)
            ts.append(time.perf_counter() - c0)
        bvh.close()
        nrays = n + int(sh_np["valid"].sum())
        cpu = {"value": round(nrays / statistics.median(ts) / 1e6, 3), "unit": "Mrays/s",
               "cores": thr, "kind": "port", "cpu_model": cpu_model(),
               "sample": "full workload (primary + shadow rays), median of 5 runs"}
    result["roofline"] = roofline
    result["cpu_baseline"] = cpu
    sc.close()
    ctx.close()
    return result


def bench_serving(args, rt, scenes, torch, rates, seed0=200):
    """Serving form: 16 resident cluster scenes (triangle counts log-uniform in [1e3, 1e6]),
    Poisson arrivals at each rate for about `serve_seconds` (at least 200, at most 2,000
    requests), resolutions uniform in {128, 256, 512, 1024} squared, 1-4 spp, primary + shadow
    rays.  p50 / p99 latency and aggregate Mrays/s per rate; sampled requests are traced again
    alone and must give the same bits as in the mix."""
    ctx = rt.Context(0)
    log("serving: building 16 resident scenes")
    scs = []
    for i in range(16):
        g = np.random.Generator(np.random.PCG64(seed0 + i))
        ntri = int(round(np.exp(g.uniform(np.log(1e3), np.log(1e6)))))
        scs.append(rt.Scene(ctx, scenes.cluster_scene(ntri, seed=seed0 + i), desc="rayhierarchy"))
    srv = rt.Server(ctx, scs, max_active=256)
    points = []
    samples = []  # (request, result in the mix)
    try:
        for rate in rates:
            g = np.random.Generator(np.random.PCG64(seed0 + int(rate)))
            nreq = int(min(2000, max(200, rate * args.serve_seconds)))
            gaps = g.exponential(1.0 / rate, size=nreq)
            reqs = [dict(scene=int(g.integers(0, 16)), width=int(s), height=int(s),
                         spp=int(g.integers(1, 5)), origin=(0.0, 0.0, -12.0), fov_deg=50.0,
                         light=(0.0, 8.0, -8.0), seed=k)
                    for k, s in enumerate(g.choice([128, 256, 512, 1024], size=nreq))]
            log(f"serving: {nreq} requests at {rate:g} req/s")
            t_start = time.perf_counter()
            due = t_start
            tickets = []
            for k, (q, gap) in enumerate(zip(reqs, gaps)):
                due += gap
                while True:
                    left = due - time.perf_counter()
                    if left <= 0:
                        break
                    time.sleep(min(left, 0.002))
                keep = k % max(1, nreq // 8) == 0 and q["width"] <= 512
                tickets.append((srv.submit(**q, outputs=keep), keep, q))
            res = []
            for tk, keep, q in tickets:
                r = srv.wait(tk, timeout_s=600.0)
                res.append(r)
                if keep:
                    samples.append((q, r))
            t_end = time.perf_counter()
            lat = np.array([r["latency_s"] for r in res]) * 1e3
            rays = sum(int(r["primary_rays"]) + int(r["shadow_rays"]) for r in res)
            first = min(r["submit_s"] for r in res)
            last = max(r["done_s"] for r in res)
            points.append({"rate_req_s": rate, "requests": nreq,
                           "p50_ms": round(float(np.percentile(lat, 50)), 2),
                           "p99_ms": round(float(np.percentile(lat, 99)), 2),
                           "mrays_s": round(rays / (last - first) / 1e6, 1),
                           "wall_s": round(t_end - t_start, 2)})
        ticks, most = srv.counters()
        # the sampled requests again, each alone: identical bits required
        log(f"serving: {len(samples)} sampled requests alone")
        mism = 0
        for q, r in samples:
            a = srv.wait(srv.submit(**q, outputs=True), timeout_s=600.0)
            mism += int((a["id"] != r["id"]).sum()) + int((a["occ"] != r["occ"]).sum())
            mism += int((a["t"].view(np.uint32) != r["t"].view(np.uint32)).sum())
    finally:
        srv.close()
        for sc in scs:
            sc.close()
        ctx.close()
    return {"workload": "serving_16_scenes", "points": points, "ticks": ticks,
            "largest_tick_jobs": most,
            "alone_vs_mix": {"requests": len(samples), "mismatches": mism},
            "rays_counted": "camera rays + valid shadow rays"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="uneven_batch_64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="uneven_batch_64: keep the built-in launch shape")
    ap.add_argument("--oracle-scenes", type=int, default=8,
                    help="uneven_batch_64: scenes whose sampled blocks are checked by the oracle")
    ap.add_argument("--extra", default="torus_rayhierarchy_512",
                    help="plain workloads reported beside the default line (comma list; "
                         "empty for none)")
    ap.add_argument("--serve-rates", default="10,20,50,100,200,500",
                    help="request rates (req/s) of the serving leg; empty to skip it")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="arrival time per rate (at least 200, at most 2,000 requests)")
    ap.add_argument("--ray-closest", type=int, default=0,
                    help="plain workloads: per-ray threshold of closest-hit traces (0: built-in)")
    ap.add_argument("--ray-occluded", type=int, default=0,
                    help="plain workloads: per-ray threshold of occlusion traces (0: built-in)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the run, one plain device copy of this many MB (calibrates "
                         "profiler byte counters against a kernel of known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding the "
              "jobs over several GPUs is not launched by this benchmark.", file=sys.stderr)
        sys.exit(2)

    import torch

    rt = _load("pbrt_v2_amd", ROOT / "pbrt-v2_amd" / "__init__.py", ROOT / "pbrt-v2_amd")
    scenes = _load("pbrt_v2_amd.scenes", ROOT / "pbrt-v2_amd" / "scenes.py")
    orc = _load("rt_oracle_py", ROOT / "oracle" / "oracle.py")
    torch.cuda.set_device(0)
    if args.workload == "uneven_batch_64":
        result = bench_uneven(args, rt, scenes, torch, orc)
        extras = [x for x in args.extra.split(",") if x]
        for x in extras:
            sub = bench_plain(args, rt, scenes, torch, orc, x)
            result[x] = {k: sub[k] for k in ("value", "unit", "ms_per_step", "config",
                                             "kernels_ms", "parity", "roofline", "cpu_baseline")
                         if k in sub}
    else:
        wl = scenes.WORKLOADS[args.workload]
        if "tiles" in wl:
            bench_frame(args, rt, scenes, torch, orc, wl)
            return
        result = bench_plain(args, rt, scenes, torch, orc, args.workload)
    if args.copy_mb > 0:
        src = torch.empty(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.serve_rates:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        result["serving"] = bench_serving(args, rt, scenes, torch, rates)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
