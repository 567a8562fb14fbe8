"""GPU: the batched form (rt_batch_*) against the plain per-scene calls and the oracle.

The batched form must give, bit for bit, the results of running every job alone: the same
hit ids, t, u, v and occlusion flags.  Covered: scenes of different sizes and node types in
one batch, closest primary / secondary and occlusion jobs mixed, jobs of the 256-ray and the
1024-ray block kernels in one group, launch shapes (grid from one work-group to more than
the resident slots, several unit sizes), HIP-graph replay (several replays with eager work in
between, re-capture after the jobs change, and a batch's groups as parallel graph branches),
and the autotuner (which must not change any result).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(h):
    return dict(id=h.id.cpu().numpy(), t=h.t.cpu().numpy(), u=h.u.cpu().numpy(),
                v=h.v.cpu().numpy())


def _same(a, b, label):
    for k in ("id", "t", "u", "v"):
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        mism = int((x != y).sum())
        assert mism == 0, f"{label}: {mism} {k} mismatches"


def _jobs(rt, scenes, gpu_ctx):
    """(scene, rays dict, mode) jobs: torus window box_dir, random1k box_dir and ia, a
    cluster scene with an unstructured ray set, a height-4 hierarchy with 2x4 leaves (512-ray
    blocks, the large-block kernel)."""
    torus = scenes.torus_scene()
    rnd = scenes.random_triangles(1024, seed=1)
    clus = scenes.cluster_scene(3000, seed=7)
    sc = {
        "torus": rt.Scene(gpu_ctx, torus, desc="rayhierarchy"),
        "rnd": rt.Scene(gpu_ctx, rnd, desc="rayhierarchy"),
        "rnd_ia": rt.Scene(gpu_ctx, rnd, desc=rt.default_desc("rayhierarchy", node="ia")),
        "clus": rt.Scene(gpu_ctx, clus, desc="rayhierarchy"),
        "h4": rt.Scene(gpu_ctx, rnd, desc=rt.default_desc("rayhierarchy", chunk_x=2, chunk_y=4,
                                                          height=4)),
    }
    full = scenes.pinhole_rays(512, 512)
    win = scenes.subimage(full, 192, 256, 128, 128)
    win["width"], win["height"] = 128, 128
    cam = scenes.pinhole_rays(96, 64)
    rr = scenes.random_rays(64 * 64, seed=5, lo=-1.5, hi=1.5, maxt=np.inf, width=64)
    rr["valid"] = (np.random.default_rng(9).random(64 * 64) < 0.75).astype(np.uint8)
    jobs = [
        ("torus", win, 0), ("rnd", cam, 0), ("rnd_ia", cam, 0), ("clus", rr, 1),
        ("h4", scenes.pinhole_rays(128, 64), 0), ("rnd", rr, 1),
    ]
    return sc, jobs


def _shadow_of(scenes, rays, ref, light=(2.0, 2.0, -2.0)):
    sh = scenes.shadow_rays_np(rays, ref["id"], ref["t"], light)
    sh["width"], sh["height"] = rays["width"], rays["height"]
    return sh


def test_batch_equals_plain_and_oracle(rt, scenes, orc, gpu_ctx):
    sc, jobs = _jobs(rt, scenes, gpu_ctx)
    plain, dev = [], []
    for name, rays, mode in jobs:
        d = rt.DeviceRays.from_numpy(rays)
        dev.append(d)
        plain.append(_np(sc[name].trace_closest(d, mode)))
    # shadow jobs from the plain hits
    shadows = [_shadow_of(scenes, rays, plain[i]) for i, (_, rays, _) in enumerate(jobs)]
    dsh = [rt.DeviceRays.from_numpy(s) for s in shadows]
    plain_occ = [sc[name].trace_occluded(dsh[i]).cpu().numpy()
                 for i, (name, _, _) in enumerate(jobs)]

    b = rt.Batch(gpu_ctx)
    outs, occs = [], []
    for i, (name, rays, mode) in enumerate(jobs):
        outs.append(b.add_closest(sc[name], dev[i], mode))
        occs.append(b.add_occluded(sc[name], dsh[i]))
    print(f"[batch] {len(jobs)} closest + {len(jobs)} occlusion jobs", flush=True)
    b.run()
    for i, (name, rays, mode) in enumerate(jobs):
        _same(_np(outs[i]), plain[i], f"job {i} {name} mode {mode}")
        assert np.array_equal(occs[i].cpu().numpy(), plain_occ[i]), f"job {i} occlusion"
    # oracle on the small jobs (the torus window and the cluster scene are covered by the
    # plain-path parity tests)
    for i in (1, 2, 4, 5):
        name, rays, mode = jobs[i]
        p = orc.params("rayhierarchy", node="ia" if name == "rnd_ia" else "box_dir",
                       cx=2 if name == "h4" else 4, height=4 if name == "h4" else 3)
        ref = orc.trace(sc[name].v9.reshape(-1, 9), rays, p, mode)
        assert np.array_equal(plain[i]["id"], ref["id"]), f"job {i} vs oracle"
        oref = orc.trace(sc[name].v9.reshape(-1, 9), shadows[i], p, orc.OCCLUDED)["occ"]
        assert np.array_equal(plain_occ[i], oref), f"job {i} occlusion vs oracle"
    # run again: identical
    b.run()
    for i in range(len(jobs)):
        _same(_np(outs[i]), plain[i], f"rerun job {i}")
    b.close()
    for s in sc.values():
        s.close()


def test_batch_launch_shapes_same_results(rt, scenes, gpu_ctx):
    """The persistent grid and the unit size change no result: the default (one work-group per
    resident slot), a single work-group taking every unit of the group, small odd grids and
    more work-groups than slots, each with several unit sizes."""
    sc, jobs = _jobs(rt, scenes, gpu_ctx)
    b = rt.Batch(gpu_ctx)
    outs, occs, dev = [], [], []
    for name, rays, mode in jobs:
        d = rt.DeviceRays.from_numpy(rays)
        dev.append(d)
        outs.append(b.add_closest(sc[name], d, mode))
    b.run()
    ref = [_np(o) for o in outs]
    shadows = [rt.DeviceRays.from_numpy(_shadow_of(scenes, rays, ref[i]))
               for i, (_, rays, _) in enumerate(jobs)]
    for i, (name, _, _) in enumerate(jobs):
        occs.append(b.add_occluded(sc[name], shadows[i]))
    b.run()
    ref_occ = [o.cpu().numpy() for o in occs]
    for grid in (1, 3, 257, 4096, 0):
        for per in (1, 5, 0):
            for p_ in (rt.RT_PARAM_GRID_CLOSEST, rt.RT_PARAM_GRID_OCCLUDED):
                b.set_param(p_, grid)
            for p_ in (rt.RT_PARAM_CLUSTERS_CLOSEST, rt.RT_PARAM_CLUSTERS_OCCLUDED):
                b.set_param(p_, per)
            print(f"[batch] grid {grid} clusters/unit {per}", flush=True)
            for o in outs:
                o.id.fill_(-7)
            for o in occs:
                o.fill_(7)
            b.run()
            for i in range(len(jobs)):
                _same(_np(outs[i]), ref[i], f"grid {grid} per {per} job {i}")
                assert np.array_equal(occs[i].cpu().numpy(), ref_occ[i]), \
                    f"grid {grid} per {per} job {i} occlusion"
    # the 1024-ray LDS tile for every block of the block path
    for tile in (1, 0):
        for p_ in (rt.RT_PARAM_TILE_CLOSEST, rt.RT_PARAM_TILE_OCCLUDED):
            b.set_param(p_, tile)
        print(f"[batch] tile {tile}", flush=True)
        for o in outs:
            o.id.fill_(-7)
        for o in occs:
            o.fill_(7)
        b.run()
        for i in range(len(jobs)):
            _same(_np(outs[i]), ref[i], f"tile {tile} job {i}")
            assert np.array_equal(occs[i].cpu().numpy(), ref_occ[i]), f"tile {tile} job {i} occ"
    with pytest.raises(rt.RtError):
        b.set_param(rt.RT_PARAM_TILE_CLOSEST, 2)
    # the per-ray path (every block) with each walk length and per-ray grid
    for p_ in (rt.RT_PARAM_RAY_CLUSTERS_CLOSEST, rt.RT_PARAM_RAY_CLUSTERS_OCCLUDED):
        b.set_param(p_, 1)
    for walk in (1, 2, 3, 7):
        for rgrid in (1, 7, 0):
            for p_ in (rt.RT_PARAM_WALK_CLOSEST, rt.RT_PARAM_WALK_OCCLUDED):
                b.set_param(p_, walk)
            for p_ in (rt.RT_PARAM_RAY_GRID_CLOSEST, rt.RT_PARAM_RAY_GRID_OCCLUDED):
                b.set_param(p_, rgrid)
            print(f"[batch] per-ray walk {walk} grid {rgrid}", flush=True)
            for o in outs:
                o.id.fill_(-7)
            for o in occs:
                o.fill_(7)
            b.run()
            for i in range(len(jobs)):
                _same(_np(outs[i]), ref[i], f"walk {walk} rgrid {rgrid} job {i}")
                assert np.array_equal(occs[i].cpu().numpy(), ref_occ[i]), \
                    f"walk {walk} rgrid {rgrid} job {i} occlusion"
    assert b.get_param(rt.RT_PARAM_WALK_CLOSEST) == 7
    b.set_param(rt.RT_PARAM_WALK_CLOSEST, 0)
    assert b.get_param(rt.RT_PARAM_WALK_CLOSEST) == 3
    with pytest.raises(rt.RtError):
        b.set_param(rt.RT_PARAM_WALK_OCCLUDED, 8)
    b.close()
    for s_ in sc.values():
        s_.close()


def test_batch_graph_replay_and_autotune(rt, scenes, gpu_ctx):
    import torch

    sc, jobs = _jobs(rt, scenes, gpu_ctx)
    dev = [rt.DeviceRays.from_numpy(r) for _, r, _ in jobs]
    ref = [_np(sc[n].trace_closest(dev[i], m)) for i, (n, _, m) in enumerate(jobs)]
    stream = torch.cuda.Stream()
    b = rt.Batch(gpu_ctx)
    b.set_param(rt.RT_PARAM_GRAPH, 1)
    outs = [b.add_closest(sc[n], dev[i], m) for i, (n, _, m) in enumerate(jobs)]
    for rep in range(4):
        for o in outs:  # poison the outputs between replays
            o.id.fill_(-7)
            o.t.fill_(-1.0)
        torch.cuda.synchronize()
        b.run(stream=stream)
        stream.synchronize()
        for i in range(len(jobs)):
            _same(_np(outs[i]), ref[i], f"graph replay {rep} job {i}")
        # eager work between replays
        sc["rnd"].trace_closest(dev[1], 0)
        torch.cuda.synchronize()
    print("[batch] graph replays ok", flush=True)
    # change the job list: re-capture
    b.clear()
    outs = [b.add_closest(sc[n], dev[i], m) for i, (n, _, m) in enumerate(jobs) if i % 2 == 0]
    b.run(stream=stream)
    stream.synchronize()
    for k, i in enumerate(range(0, len(jobs), 2)):
        _same(_np(outs[k]), ref[i], f"re-captured job {i}")
    # autotune, then results unchanged
    b.set_param(rt.RT_PARAM_GRAPH, 0)
    b.autotune(reps=1, stream=stream)
    stream.synchronize()
    tuned = [b.get_param(p) for p in (rt.RT_PARAM_CLUSTERS_CLOSEST, rt.RT_PARAM_GRID_CLOSEST)]
    print(f"[batch] tuned clusters/unit, grid = {tuned}", flush=True)
    for o in outs:
        o.id.fill_(-7)
    b.run(stream=stream)
    stream.synchronize()
    for k, i in enumerate(range(0, len(jobs), 2)):
        _same(_np(outs[k]), ref[i], f"tuned job {i}")
    b.close()
    for s in sc.values():
        s.close()


def test_batch_raybvh_task_jobs(rt, scenes, gpu_ctx):
    """ray-bvh jobs in a batch, as a renderer's tasks of one frame: four task windows of one
    scene (camera rays, then bounce rays from their hits) and their shadow rays; every job
    bit-identical to its plain call.  The plain calls are pinned against the oracle by the
    parity tests."""
    v9 = scenes.room_scene(nclusters=40)
    sc = rt.Scene(gpu_ctx, v9, desc="ray-bvh")
    full = scenes.pinhole_rays(128, 128, origin=(0.0, 0.0, -9.0), fov_deg=50.0)
    wins = []
    for y0 in (0, 64):
        for x0 in (0, 64):
            w = scenes.subimage(full, x0, y0, 64, 64)
            w["width"], w["height"] = 64, 64
            wins.append(rt.DeviceRays.from_numpy(w))
    plain = [sc.trace_closest(d, rt.RT_CLOSEST_PRIMARY) for d in wins]
    bounce = [rt.make_bounce_rays(sc, d, h, seed=5 + k, depth=1)
              for k, (d, h) in enumerate(zip(wins, plain))]
    plain_b = [sc.trace_closest(d, rt.RT_CLOSEST_SECONDARY) for d in bounce]
    shadow = [rt.make_shadow_rays(d, h, (1.0, 2.5, -8.0)) for d, h in zip(wins, plain)]
    plain_o = [sc.trace_occluded(d) for d in shadow]
    b = rt.Batch(gpu_ctx)
    outs = [b.add_closest(sc, d, rt.RT_CLOSEST_PRIMARY) for d in wins]
    outs_b = [b.add_closest(sc, d, rt.RT_CLOSEST_SECONDARY) for d in bounce]
    occs = [b.add_occluded(sc, d) for d in shadow]
    b.set_param(rt.RT_PARAM_GRAPH, 1)
    print(f"[batch] ray-bvh: {len(outs) + len(outs_b)} closest + {len(occs)} occlusion jobs",
          flush=True)
    import torch

    stream = torch.cuda.Stream()  # graph capture needs a stream of its own
    for rep in range(3):  # capture and launch, then graph replays (three groups as branches)
        for o in outs + outs_b:
            o.id.fill_(-7)
        for o in occs:
            o.fill_(7)
        torch.cuda.synchronize()
        b.run(stream=stream)
        stream.synchronize()
        for k in range(len(wins)):
            _same(_np(outs[k]), _np(plain[k]), f"rep {rep} task {k} primary")
            _same(_np(outs_b[k]), _np(plain_b[k]), f"rep {rep} task {k} bounce")
            assert bool((occs[k] == plain_o[k]).all()), f"rep {rep} task {k} occlusion"
    assert sum(int((h.id >= 0).sum().item()) for h in plain_b) > 0
    b.close()
    sc.close()


def test_batch_graph_branches_closest_and_occlusion(rt, scenes, gpu_ctx):
    """A captured batch whose groups (closest hits by node type and mode, and occlusion) run
    as parallel graph branches on side streams: four replays, outputs poisoned before each and
    eager work between them, every job bit-identical to its plain call each time."""
    import torch

    sc, jobs = _jobs(rt, scenes, gpu_ctx)
    dev = [rt.DeviceRays.from_numpy(r) for _, r, _ in jobs]
    plain = [sc[n].trace_closest(dev[i], m) for i, (n, _, m) in enumerate(jobs)]
    ref = [_np(h) for h in plain]
    prim = [i for i, (_, _, m) in enumerate(jobs) if m == 0]
    shadow = [rt.make_shadow_rays(dev[i], plain[i], (2.0, 2.0, -2.0)) for i in prim]
    ref_occ = [sc[jobs[i][0]].trace_occluded(s).cpu().numpy() for i, s in zip(prim, shadow)]
    stream = torch.cuda.Stream()
    b = rt.Batch(gpu_ctx)
    b.set_param(rt.RT_PARAM_GRAPH, 1)
    outs = [b.add_closest(sc[n], dev[i], m) for i, (n, _, m) in enumerate(jobs)]
    occs = [b.add_occluded(sc[jobs[i][0]], s) for i, s in zip(prim, shadow)]
    print(f"[batch] graph branches: {len(outs)} closest + {len(occs)} occlusion jobs",
          flush=True)
    for rep in range(4):
        for o in outs:
            o.id.fill_(-7)
            o.t.fill_(-1.0)
        for o in occs:
            o.fill_(7)
        torch.cuda.synchronize()
        b.run(stream=stream)
        stream.synchronize()
        for i in range(len(jobs)):
            _same(_np(outs[i]), ref[i], f"replay {rep} job {i}")
        for k in range(len(occs)):
            assert np.array_equal(occs[k].cpu().numpy(), ref_occ[k]), \
                f"replay {rep} occlusion job {k}"
        sc["rnd"].trace_occluded(shadow[1])  # eager work between replays
        torch.cuda.synchronize()
    b.close()
    for s in sc.values():
        s.close()
