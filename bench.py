#!/usr/bin/env python3
"""Throughput of the strand-physics step on one MI355X.

Workload (default): one groom of 100,000 strands x 32 points fp32, roots on the
unit upper hemisphere, lengths U[0.8,1.2], sinusoidal x wind (amplitude 2,
period 1 s) + per-strand N(0, 0.1^2) noise, sphere r 0.3 orbiting at radius
1.1, hair-hair friction 0.05 through a 64^3 voxel grid over [-2.2,2.2]^3,
dt 1/60.  A "step" is one full physics step of the groom (predict, FTL, DFTL,
collider penalty, grid splat, grid friction).  State is resident in HBM before
the timed region starts.

The scene is 240 steps (step k sees totalTime (k+1) dt); --warmup W (default
10) steps run untimed, then --steps K (default: the rest of the scene) are
timed, so by default the timed steps are the scene's steps 10..239.

Prints ONE JSON line.  `value` is point-updates/s (S*N per step) over the
timed steps (wall clock between two stream synchronisations).  `roofline`
prices the dominant kernel, the strand step kernel, by its algorithmic bytes
(48 B per non-root point + 16 B per strand) over its average duration,
measured with HIP events the library records on the stream the kernel runs
on, over the same scene steps replayed from a fresh copy of the groom (so
the kernel averages describe exactly the timed steps; the replay must end
bit-identical to the timed run); the voxel reduce kernel's share (32 B per
voxel) and the whole step's figure are given beside it.  `traffic` is read
from a committed rocprofv3 counter summary, named in `traffic_source`.
`cpu_baseline` times the oracle's FP32 restatement of the reference shader on
the host cores on a bounded sample of the same workload.

--workload batched-256: 256 grooms of uneven shapes stepped together with one
captured graph launch per step, reported beside the same grooms stepped with
one launch pair per groom (`per_groom_launch`).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "realtime-vulkan-hair_amd"))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0


def make_params(hairsim, wl, k, mesh_id=0):
    return hairsim.step_params(wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                               friction=wl.friction, num_colliders=1, hair_hair=wl.hair_hair, mesh_id=mesh_id)


def cpu_baseline(wl, budget_s=15.0):
    """Time the FP32 oracle (OpenMP over strands) on a bounded sample of the
    workload: every strand (a strand subset with a mesh collider, which the
    oracle checks by scanning every triangle), as many steps as fit in
    budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the oracle, used here only as the timed CPU baseline
    from hairsim import scenes

    g = wl.grooms[0]
    # the oracle scans every mesh triangle per point: a smaller sample then
    S = min(g.S, 2000 if wl.mesh is not None else 1_000_000)
    strands = g.strands[:S].copy()
    seg = g.seg[:S].copy()
    noise = None if g.noise is None else g.noise[:S].copy()
    threads = min(os.cpu_count() or 1, 16)
    grid = None
    recs = None if wl.mesh is None else oracle.mesh_prepare(*wl.mesh)
    steps = 0
    t0 = time.perf_counter()
    while True:
        p = oracle.Params(dt=wl.dt, wind=wl.wind(steps), damping=wl.damping, sphere_radius=wl.sphere_radius,
                          friction=wl.friction, hair_hair=wl.hair_hair, grid_dim=g.grid_dim,
                          grid_origin=g.grid_origin, grid_h=g.grid_h,
                          colliders=scenes.sphere_collider(wl.sphere(steps)), seg_length=seg, wind_noise=noise,
                          mesh=recs, mesh_margin=wl.mesh_margin)
        grid = oracle.step32_mt(p, strands, g.N, grid, threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= wl.steps:
            break
    return {"value": S * g.N * steps / el, "unit": "point-updates/s", "cores": threads, "kind": "port",
            "sample": f"{S} of {g.S} strands x {g.N} points, {g.grid_dim}^3 grid, {steps} steps in {el:.1f} s "
                      f"(FP32 C restatement of compute.comp, OpenMP"
                      f"{'; mesh contact by scanning every triangle' if recs is not None else ''})"}


def copy_probe(mib, reps=5):
    """A plain elementwise kernel of known traffic (read n bytes, write n bytes):
    the calibration that scales the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "reps": reps}),
          flush=True)


def load_traffic(name):
    """HBM bytes per dispatch of a kernel and where the figure comes from.

    Counters cannot be read from inside this process: the figure is taken from
    a committed summary of two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE) of the command named there, scaled by the calibration factors
    it lists.  (None, None) when no summary exists."""
    prof = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(prof):
        return None, None
    with open(prof) as f:
        t = json.load(f)
    src = (f"profiles/traffic_{name}.json (committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs"
           f"{' of: ' + t['command'] if t.get('command') else ''}"
           f"{', dispatches ' + t['dispatches'] if t.get('dispatches') else ''}; not measured in this run)")
    return t.get("hbm_bytes_per_launch"), src


def scene_steps(args, wl):
    """(warmup, steps) for a single-groom workload: by default the whole
    scene, its first 10 steps untimed."""
    w = 10 if args.warmup is None else args.warmup
    k = max(1, wl.steps - w) if args.steps is None else args.steps
    return w, k


def run_scene(ctx, hairsim, scenes, wl, mesh_id, warmup, steps, profile):
    """Steps 0..warmup-1 untimed, then steps warmup..warmup+steps-1 between two
    stream synchronisations, on a fresh copy of the groom.  Scene step k sees
    totalTime (k+1)*dt and the wind and sphere of that time.  With `profile`
    the library brackets each kernel of the timed steps with HIP events.
    Returns (wall seconds of the timed steps, kernel ms dict, final state,
    near-limit count)."""
    g = wl.grooms[0]
    groom = ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                      grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h)
    total = warmup + steps
    # per-step inputs prepared ahead so the timed loop is launches only
    cols = [[hairsim.HsCollider.from_buffer_copy(scenes.sphere_collider(wl.sphere(k)).tobytes())]
            for k in range(total)]
    prm = [make_params(hairsim, wl, k, mesh_id) for k in range(total)]
    tms = [float(wl.time(k)) for k in range(total)]
    for k in range(warmup):
        groom.step(wl.dt, tms[k], cols[k], prm[k])
    ctx.sync()
    if profile:
        ctx.profile_enable(True)
    t0 = time.perf_counter()
    for k in range(warmup, total):
        groom.step(wl.dt, tms[k], cols[k], prm[k])
    ctx.sync()
    el = time.perf_counter() - t0
    kern = None
    if profile:
        n_launch, _ = ctx.profile_read()
        step_ms, reduce_ms = ctx.profile_breakdown()
        kern = {"launches": n_launch, "step_ms": step_ms, "reduce_ms": reduce_ms, "mesh_ms": ctx.profile_mesh()}
        ctx.profile_enable(False)
    out, _ = groom.read()
    near = groom.grid_near_limit() if g.grid_dim and wl.hair_hair else 0
    groom.destroy()
    return el, kern, out, near


def tune_tile(ctx, hairsim, scenes, wl, mesh_id, warmup):
    """The tile (shape, LDS windows, workgroup order) for the timed steps:
    hs_autotune_batch on copies of a groom stepped through the untimed steps,
    each candidate timed over the first timed steps' inputs."""
    g = wl.grooms[0]
    groom = ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                      grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h)
    for k in range(warmup):
        groom.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], make_params(hairsim, wl, k, mesh_id))
    b = ctx.batch([groom])
    b._params[0] = make_params(hairsim, wl, warmup, mesh_id)
    b.set_colliders(scenes.sphere_collider(wl.sphere(warmup))[None, None])
    b.set_time(wl.dt, float(wl.time(warmup)))
    tile, ms = ctx.autotune(b, steps=5)
    ctx.sync()
    groom.destroy()
    return tile, ms


def tile_record(tile, tuned_ms):
    return {"gather_cells": tile[0], "splat_cells": tile[1], "xcd_order": tile[2], "strands_per_workgroup": tile[3],
            "points_per_segment": tile[4], "strands_per_wavefront": 64, "autotuned_ms_per_step": tuned_ms}


def bench_single(args, hairsim, scenes, wl):
    g = wl.grooms[0]
    warmup, steps = scene_steps(args, wl)
    args.warmup, args.steps = warmup, steps
    ctx = hairsim.Context(0)
    mesh_id = 0
    if wl.mesh is not None:
        mesh_id = ctx.mesh(*wl.mesh, grid_dim=wl.mesh_grid, margin=wl.mesh_margin)
    if args.tile:
        ctx.set_tile(*args.tile)
        tile, tuned_ms = ctx.get_tile(), None
    elif not args.no_autotune:
        tile, tuned_ms = tune_tile(ctx, hairsim, scenes, wl, mesh_id, warmup)
    else:
        tile, tuned_ms = ctx.get_tile(), None
    # the timed run, then the same scene steps again from a fresh copy of the
    # groom with kernel events: the kernel averages describe the timed steps
    el, _, out, near = run_scene(ctx, hairsim, scenes, wl, mesh_id, warmup, steps, profile=False)
    _, kern, out2, _ = run_scene(ctx, hairsim, scenes, wl, mesh_id, warmup, steps, profile=True)
    ctx.close()
    if not np.isfinite(out).all():
        print("bench.py: non-finite state after the run", file=sys.stderr)
        sys.exit(1)
    if not np.array_equal(out.view(np.uint32), out2.view(np.uint32)):
        print("bench.py: the profiled run differs from the timed run", file=sys.stderr)
        sys.exit(1)
    n_launch = kern["launches"]
    step_ms, reduce_ms, mesh_ms = kern["step_ms"], kern["reduce_ms"], kern["mesh_ms"]

    pts = g.S * g.N
    value = pts * args.steps / el
    n = max(1, n_launch)
    step_us = step_ms / n * 1e3
    reduce_us = reduce_ms / n * 1e3
    mesh_us = mesh_ms / n * 1e3
    b_step_kernel = scenes.strand_bytes(g.S, g.N)
    b_grid = scenes.algorithmic_bytes(g.S, g.N, g.grid_dim, wl.hair_hair) - b_step_kernel
    b_mesh = 0
    if wl.mesh is not None:
        # mesh pass: reads every stored position (12 B), writes its contact
        # (16 B); the step kernel reads the contact of points 1..N-1
        b_mesh = 28 * g.S * g.N
        b_step_kernel += 16 * g.S * (g.N - 1)
    achieved = b_step_kernel / (step_us * 1e-6) / 1e9
    whole = (b_step_kernel + b_grid + b_mesh) / ((step_us + reduce_us + mesh_us) * 1e-6) / 1e9
    traffic, traffic_src = load_traffic(wl.name)
    res = {
        "metric": "strand-point updates/sec",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {g.S} strands x {g.N} points, grid {g.grid_dim}^3, "
                               f"wind + noise, orbiting sphere, friction {wl.friction}"
                               + (f", head mesh {len(wl.mesh[1])} triangles" if wl.mesh is not None else ""),
                   "strands": g.S, "points": g.N, "grid_dim": g.grid_dim, "hair_hair": wl.hair_hair,
                   "scene_steps": f"{args.warmup}..{args.warmup + args.steps - 1} timed of a {wl.steps}-step scene"
                                  + (" (the forcing continues past its end)"
                                     if args.warmup + args.steps > wl.steps else "")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel_timing": "HIP events around each kernel, recorded on the library's stream over the "
                                      "same scene steps in a second run from a fresh copy of the groom",
                     "kernel": "k_step", "kernel_avg_us": step_us, "algorithmic_bytes_per_launch": b_step_kernel,
                     "reduce_kernel_avg_us": reduce_us, "reduce_algorithmic_bytes": b_grid,
                     "whole_step_achieved": whole, "whole_step_frac": whole / PEAK_HBM_GBS},
        "grid_cells_near_int32_limit": near,
        "tile": tile_record(tile, tuned_ms),
    }
    if b_mesh:
        mesh_gbs = b_mesh / (mesh_us * 1e-6) / 1e9
        mesh = {"kernel": "k_mesh_force", "kernel_avg_us": mesh_us, "algorithmic_bytes_per_launch": b_mesh,
                "achieved": mesh_gbs, "frac": mesh_gbs / PEAK_HBM_GBS}
        if mesh_us > step_us:
            # the mesh pass is the dominant kernel: it heads the roofline
            rl = res["roofline"]
            m_traffic, m_src = load_traffic(wl.name + "-mesh")
            res["roofline"] = {"bound": "hbm", "achieved": mesh_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": mesh_gbs / PEAK_HBM_GBS, "traffic": m_traffic, "traffic_source": m_src,
                               "kernel_timing": rl["kernel_timing"],
                               "kernel": "k_mesh_force", "kernel_avg_us": mesh_us,
                               "algorithmic_bytes_per_launch": b_mesh,
                               "step_kernel": {k: rl[k] for k in ("achieved", "frac", "kernel_avg_us",
                                                                  "algorithmic_bytes_per_launch", "traffic",
                                                                  "traffic_source")},
                               "reduce_kernel_avg_us": reduce_us, "reduce_algorithmic_bytes": b_grid,
                               "whole_step_achieved": whole, "whole_step_frac": whole / PEAK_HBM_GBS}
        else:
            res["roofline"]["mesh_pass"] = mesh
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(res), flush=True)


def bench_batched(args, hairsim, scenes, wl):
    """Many grooms per step: hs_step_batch with captured graphs, against the
    same grooms stepped with one hs_step launch pair per groom."""
    ctx = hairsim.Context(0)
    grooms = [ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                        grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h) for g in wl.grooms]
    n = len(grooms)
    ctx.set_batch_graphs(not args.no_graphs)
    b = ctx.batch(grooms, max_colliders=1)
    b.params["damping"] = wl.damping
    b.params["sphere_radius"] = wl.sphere_radius
    b.params["friction"] = wl.friction
    b.params["hair_hair"] = 1 if wl.hair_hair else 0
    b.params["wind"] = np.stack([wl.wind(0, gi) for gi in range(n)])
    b.set_colliders(np.stack([scenes.sphere_collider(wl.sphere(0, gi)) for gi in range(n)])[:, None])

    def batch_step(k):
        b.set_time(wl.dt, float(wl.time(k)))
        b.step()

    args.warmup, args.steps = scene_steps(args, wl)
    for k in range(args.warmup):
        batch_step(k)
    # tile choice: timed on copies of the grooms at the first timed step
    # (their state is untouched)
    b.set_time(wl.dt, float(wl.time(args.warmup)))
    if args.tile:
        ctx.set_tile(*args.tile)
        tile, tuned_ms = ctx.get_tile(), None
    elif not args.no_autotune:
        tile, tuned_ms = ctx.autotune(b, steps=5)
    else:
        tile, tuned_ms = ctx.get_tile(), None
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        batch_step(k)
    ctx.sync()
    el = time.perf_counter() - t0
    # per-kernel durations: the grooms restored to their initial state, the
    # same scene steps again with eager launches bracketed by HIP events
    for gr, g in zip(grooms, wl.grooms):
        gr.write(g.strands)
    ctx.set_batch_graphs(False)
    for k in range(args.warmup):
        batch_step(k)
    ctx.sync()
    ctx.profile_enable(True)
    prof_steps = args.steps
    for k in range(args.warmup, args.warmup + args.steps):
        batch_step(k)
    ctx.sync()
    step_ms, reduce_ms = ctx.profile_breakdown()
    ctx.profile_enable(False)

    # the plain form: one hs_step per groom per step (inputs prebuilt)
    lib = hairsim.lib()
    per = []
    for gi, gr in enumerate(grooms):
        t = hairsim.HsTime(wl.dt, 0.0)
        cl = (hairsim.HsCollider * 1)(hairsim.HsCollider.from_buffer_copy(
            scenes.sphere_collider(wl.sphere(0, gi)).tobytes()))
        p = hairsim.step_params(wind=wl.wind(0, gi), damping=wl.damping, sphere_radius=wl.sphere_radius,
                                friction=wl.friction, num_colliders=1, hair_hair=wl.hair_hair)
        per.append((gr.id, t, cl, p))
    import ctypes as C
    plain_steps = max(3, min(args.steps // 4, 30))
    for gid, t, cl, p in per:
        ctx._check(lib.hs_step(ctx._h, gid, C.byref(t), cl, 1, C.byref(p)))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for gid, t, cl, p in per:
            ctx._check(lib.hs_step(ctx._h, gid, C.byref(t), cl, 1, C.byref(p)))
    ctx.sync()
    el_plain = time.perf_counter() - t0
    for gr in grooms:
        gr.destroy()
    ctx.close()

    pts = sum(g.S * g.N for g in wl.grooms)
    value = pts * args.steps / el
    plain = pts * plain_steps / el_plain
    b_step_kernel = sum(scenes.strand_bytes(g.S, g.N) for g in wl.grooms)
    b_all = sum(scenes.algorithmic_bytes(g.S, g.N, g.grid_dim, wl.hair_hair) for g in wl.grooms)
    step_us = step_ms / prof_steps * 1e3
    reduce_us = reduce_ms / prof_steps * 1e3
    achieved = b_step_kernel / (step_us * 1e-6) / 1e9
    whole = b_all / ((step_us + reduce_us) * 1e-6) / 1e9
    b_traffic, b_traffic_src = load_traffic(wl.name)
    res = {
        "metric": "strand-point updates/sec",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {n} grooms, {sum(g.S for g in wl.grooms)} strands, "
                               f"{pts} points, grids {wl.grooms[0].grid_dim}^3, one batched step per frame"
                               f"{'' if args.no_graphs else ' (captured graph)'}",
                   "grooms": n, "points": pts,
                   "scene_steps": f"{args.warmup}..{args.warmup + args.steps - 1} timed of a {wl.steps}-step scene"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": b_traffic, "traffic_source": b_traffic_src,
                     "kernel_timing": "HIP events around the step and reduce launches on the library's stream, "
                                      "the same scene steps again with eager launches after restoring every "
                                      "groom's initial state",
                     "kernel": "k_step_batch", "kernel_avg_us": step_us,
                     "algorithmic_bytes_per_launch": b_step_kernel, "reduce_kernel_avg_us": reduce_us,
                     "whole_step_achieved": whole, "whole_step_frac": whole / PEAK_HBM_GBS},
        "per_groom_launch": {"value": plain, "steps": plain_steps, "ms_per_step": el_plain / plain_steps * 1e3,
                             "speedup_of_batched": value / plain},
        "tile": tile_record(tile, tuned_ms),
    }
    print(json.dumps(res), flush=True)


def bench_served(args, hairsim, scenes):
    """Served jobs: Poisson arrivals at --rate jobs/s for --seconds; strands
    log-uniform in [256, 32768], points from {16, 32, 64}, 60-600 steps, wind
    and sphere keyframes every 0.25 s; 32^3 grids; admission by a 200 GB HBM
    budget; each launch takes resident jobs round-robin up to a point cap.
    Jobs still queued when arrivals end are cancelled; admitted ones run to
    completion."""
    import threading
    rng = np.random.default_rng(scenes.SEED_BASE + 3)
    # a pool of distinct job shapes; each arrival picks one (synthetic inputs)
    pool = []
    for _ in range(48):
        S = int(round(np.exp(rng.uniform(np.log(256), np.log(32768)))))
        N = int(rng.choice([16, 32, 64]))
        roots, nrm = scenes.hemisphere_roots(rng, S)
        aos, seg = scenes.straight_strands(roots, nrm, N, rng.uniform(0.8, 1.2, S))
        pool.append((aos, N, seg))
    G = 32
    grid = dict(grid_dim=G, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / G)
    prm = hairsim.step_params(damping=0.9, sphere_radius=0.3, friction=0.05, num_colliders=1, hair_hair=True)
    budget = 200 << 30
    srv = hairsim.Server(0, hbm_budget=budget, max_points_per_step=args.max_points, use_graphs=False,
                         step_loops=args.step_loops, launch_depth=args.launch_depth, tile=args.tile)
    srv_tile = args.tile or (2048, 2048, 0, 256, 8)
    n_jobs = int(args.rate * args.seconds * 1.5) + 16
    gaps = rng.exponential(1.0 / args.rate, n_jobs)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < args.seconds]
    picks = rng.integers(0, len(pool), len(arrivals))
    steps = rng.integers(60, 601, len(arrivals))
    key_sets = []
    for j in range(64):
        nk = int(np.ceil(600 / 60 / 0.25)) + 2
        k = np.zeros((nk, 7), dtype=np.float32)
        k[:, 0] = np.arange(nk) * 0.25
        k[:, 1:4] = rng.normal(0.0, 1.5, (nk, 3))
        ang = rng.uniform(0, 2 * np.pi) + np.arange(nk) * 0.4
        k[:, 4] = 1.1 * np.cos(ang)
        k[:, 5] = rng.uniform(0.0, 1.0)
        k[:, 6] = 1.1 * np.sin(ang)
        key_sets.append(hairsim.keyframes(k))
    checked = []
    t0 = time.perf_counter()
    for i, (ta, pi, ns) in enumerate(zip(arrivals, picks, steps)):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        aos, N, seg = pool[pi]
        keep = i % 500 == 0
        jid, res = srv.submit(aos, N, int(ns), 1.0 / 60.0, key_sets[i % len(key_sets)], prm, seg_length=seg,
                              keep_result=keep, **grid)
        if keep:
            checked.append((jid, res))
    t_arrivals = time.perf_counter() - t0
    cancelled = srv.cancel_pending()
    srv.drain()
    wall = time.perf_counter() - t0
    st = srv.stats()
    for jid, res in checked:
        try:
            out = srv.wait(jid, 0.0)
            if not np.isfinite(out).all():
                print("bench.py: non-finite served result", file=sys.stderr)
                sys.exit(1)
        except hairsim.HairSimError:
            pass  # cancelled before admission
    srv.close()

    # the plain form on the same shapes: one hs_step launch pair per groom
    ctx = hairsim.Context(0)
    import ctypes as C
    lib = hairsim.lib()
    plain_pts = 0
    plain_t = 0.0
    for aos, N, seg in pool[:16]:
        gr = ctx.groom(aos, N, seg_length=seg, **grid)
        t = hairsim.HsTime(1.0 / 60.0, 1.0 / 60.0)
        cl = (hairsim.HsCollider * 1)(hairsim.HsCollider.from_buffer_copy(
            scenes.sphere_collider((1.1, 0.5, 0.0)).tobytes()))
        for _ in range(3):
            ctx._check(lib.hs_step(ctx._h, gr.id, C.byref(t), cl, 1, C.byref(prm)))
        ctx.sync()
        a = time.perf_counter()
        for _ in range(20):
            ctx._check(lib.hs_step(ctx._h, gr.id, C.byref(t), cl, 1, C.byref(prm)))
        ctx.sync()
        plain_t += time.perf_counter() - a
        plain_pts += 20 * gr.S * N
        gr.destroy()
    ctx.close()
    plain = plain_pts / plain_t
    value = st["point_updates"] / wall
    res = {
        "metric": "strand-point updates/sec",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": int(st["launches"]),
        "warmup": 0,
        "ms_per_step": st["p50_step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"served: Poisson {args.rate:g} jobs/s for {args.seconds:g} s, S log-uniform "
                               f"[256, 32768], N in {{16,32,64}}, 60-600 steps, keyframes every 0.25 s, 32^3 grids, "
                               f"200 GB budget, <= {args.max_points} points per launch, round-robin, {args.step_loops} step loop(s) "
                               f"x {args.launch_depth} queued launch(es), tile {srv_tile[3]} x {srv_tile[4]}",
                   "jobs_submitted": st["submitted"], "jobs_completed": st["completed"],
                   "jobs_cancelled_unadmitted": cancelled},
        "served": {"p50_step_ms": st["p50_step_ms"], "p99_step_ms": st["p99_step_ms"],
                   "max_step_ms": st["max_step_ms"], "p50_latency_ms": st["p50_latency_ms"],
                   "p99_latency_ms": st["p99_latency_ms"], "p50_wait_ms": st["p50_wait_ms"],
                   "p99_wait_ms": st["p99_wait_ms"], "launches": st["launches"], "job_steps": st["job_steps"],
                   "peak_jobs_per_launch": st["peak_jobs_per_launch"],
                   "peak_resident_gb": st["peak_resident_bytes"] / 1e9, "busy_s": st["busy_s"],
                   "gpu_busy_s": st["gpu_busy_s"], "gpu_busy_frac": st["gpu_busy_s"] / max(st["busy_s"], 1e-9),
                   "wall_s": wall, "arrival_window_s": t_arrivals,
                   "step_ms_is": "device time of the launch a job-step ran in (HIP events)",
                   "latency_ms_is": "host enqueue of the launch to its completion seen by the host"},
        "per_groom_launch": {"value": plain, "speedup_of_served": value / plain},
    }
    print(json.dumps(res), flush=True)


def bench_export(args, hairsim, scenes, samples=64):
    """Render-facing output of the 100k x 32 groom: hs_export_strands (Strand
    AoS + indirect record into device memory) and hs_resample_strands (the
    tessellation curve, `samples` points per strand), each timed over
    args.steps back-to-back calls between two stream synchronisations."""
    wl = scenes.plain_100k()
    g = wl.grooms[0]
    ctx = hairsim.Context(0)
    groom = ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                      grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h)
    for k in range(4):  # pending grid friction: the export applies it
        groom.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], make_params(hairsim, wl, k))
    aos = hairsim.DeviceBuffer(48 * g.S * g.N)
    ind = hairsim.DeviceBuffer(16)
    crv = hairsim.DeviceBuffer(16 * g.S * samples)
    res = {}
    for name, fn, nbytes in (
            # export: pos + vel planes read, grid cells of the friction gather, AoS written
            ("hs_export_strands", lambda: groom.export(aos, ind), 24 * g.S * g.N + 48 * g.S * g.N),
            # resample: positions read once, samples written
            ("hs_resample_strands", lambda: groom.resample(samples, crv), 12 * g.S * g.N + 16 * g.S * samples)):
        for _ in range(args.warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.sync()
        us = (time.perf_counter() - t0) / args.steps * 1e6
        gbs = nbytes / (us * 1e-6) / 1e9
        res[name] = {"us_per_call": us, "algorithmic_bytes": nbytes, "achieved": gbs, "frac": gbs / PEAK_HBM_GBS}
    for b in (aos, ind, crv):
        b.free()
    groom.destroy()
    ctx.close()
    e = res["hs_export_strands"]
    print(json.dumps({
        "metric": "strand-point exports/sec", "value": g.S * g.N / (e["us_per_call"] * 1e-6),
        "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": e["us_per_call"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"export-100k: {g.S} strands x {g.N} points to the Strand layout + "
                               f"{samples} curve samples per strand, device memory"},
        "roofline": {"bound": "hbm", "achieved": e["achieved"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": e["frac"], "traffic": None, "kernel": "k_unpack"},
        "calls": res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the rest of the workload's scene after the warm-up)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before them (default 10)")
    ap.add_argument("--workload", default="plain-100k",
                    choices=["plain-100k", "large-1m", "plain-1k", "batched-256", "served", "export-100k"])
    ap.add_argument("--rate", type=float, default=2000.0, help="served: job arrivals per second")
    ap.add_argument("--seconds", type=float, default=30.0, help="served: arrival window")
    ap.add_argument("--max-points", type=int, default=8_000_000, help="served: points per batched launch")
    ap.add_argument("--step-loops", type=int, default=4, help="served: step loops (contexts with their own stream)")
    ap.add_argument("--launch-depth", type=int, default=1, help="served: launches queued per step loop")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="batched: eager launches instead of graphs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile")
    ap.add_argument("--tile", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    help="use this tile instead of autotuning: gather_cells,splat_cells,xcd_order,strands,points")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a plain elementwise kernel over MIB MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for --gpus {args.gpus}); "
              "the strand step has no multi-GPU path here", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import hairsim
    from hairsim import scenes

    if args.workload == "batched-256":
        bench_batched(args, hairsim, scenes, scenes.batched())
        return
    if args.workload == "served":
        bench_served(args, hairsim, scenes)
        return
    if args.workload == "export-100k":
        args.warmup = 5 if args.warmup is None else args.warmup
        args.steps = 20 if args.steps is None else args.steps
        bench_export(args, hairsim, scenes)
        return
    wl = {"plain-100k": scenes.plain_100k, "large-1m": scenes.large_1m, "plain-1k": scenes.plain_1k}[args.workload]()
    bench_single(args, hairsim, scenes, wl)


if __name__ == "__main__":
    main()
