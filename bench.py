#!/usr/bin/env python3
"""Throughput of the strand-physics step on one MI355X.

Workload (default): one groom of 100,000 strands x 32 points fp32, roots on the
unit upper hemisphere, lengths U[0.8,1.2], sinusoidal x wind (amplitude 2,
period 1 s) + per-strand N(0, 0.1^2) noise, sphere r 0.3 orbiting at radius
1.1, hair-hair friction 0.05 through a 64^3 voxel grid over [-2.2,2.2]^3,
dt 1/60.  A "step" is one full physics step of the groom (predict, FTL, DFTL,
collider penalty, grid splat, grid friction).  State is resident in HBM before
the timed region starts.

Prints ONE JSON line.  `value` is point-updates/s (S*N per step) over the
timed steps (wall clock between two stream synchronisations).  `roofline`
prices the step kernel's algorithmic bytes (48 B per non-root point + 16 B per
strand + 32 B per voxel) over its average duration, measured with HIP events
the library records on the stream the kernel runs on.  `cpu_baseline` times
the oracle's FP32 restatement of the reference shader on the host cores on a
bounded sample of the same workload.
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


def make_params(hairsim, wl, k):
    return hairsim.step_params(wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                               friction=wl.friction, num_colliders=1, hair_hair=wl.hair_hair)


def cpu_baseline(wl, budget_s=12.0):
    """Time the FP32 oracle (OpenMP over strands) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker, used here only as the timed CPU baseline
    from hairsim import scenes

    g = wl.grooms[0]
    S = min(g.S, 20000)
    strands = g.strands[:S].copy()
    seg = g.seg[:S].copy()
    noise = None if g.noise is None else g.noise[:S].copy()
    threads = min(os.cpu_count() or 1, 16)
    grid = None
    steps = 0
    t0 = time.perf_counter()
    while True:
        p = oracle.Params(dt=wl.dt, wind=wl.wind(steps), damping=wl.damping, sphere_radius=wl.sphere_radius,
                          friction=wl.friction, hair_hair=wl.hair_hair, grid_dim=g.grid_dim,
                          grid_origin=g.grid_origin, grid_h=g.grid_h,
                          colliders=scenes.sphere_collider(wl.sphere(steps)), seg_length=seg, wind_noise=noise)
        grid = oracle.step32_mt(p, strands, g.N, grid, threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= wl.steps:
            break
    return {"value": S * g.N * steps / el, "unit": "point-updates/s", "cores": threads, "kind": "port",
            "sample": f"{S} of {g.S} strands x {g.N} points, {g.grid_dim}^3 grid, {steps} steps in {el:.1f} s "
                      f"(FP32 C restatement of compute.comp, OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain-100k", choices=["plain-100k", "large-1m", "plain-1k"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the strand step has no multi-GPU path here", file=sys.stderr)
        sys.exit(2)

    import hairsim
    from hairsim import scenes

    wl = {"plain-100k": scenes.plain_100k, "large-1m": scenes.large_1m, "plain-1k": scenes.plain_1k}[args.workload]()
    g = wl.grooms[0]
    ctx = hairsim.Context(0)
    groom = ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                      grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h)
    # per-step inputs prepared ahead so the timed loop is launches only
    total = args.warmup + 2 * args.steps
    cols = [[hairsim.HsCollider.from_buffer_copy(scenes.sphere_collider(wl.sphere(k % wl.steps)).tobytes())]
            for k in range(total)]
    prm = [make_params(hairsim, wl, k % wl.steps) for k in range(total)]
    tms = [float(wl.time(k % wl.steps)) for k in range(total)]

    k = 0
    for _ in range(args.warmup):
        groom.step(wl.dt, tms[k], cols[k], prm[k])
        k += 1
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        groom.step(wl.dt, tms[k], cols[k], prm[k])
        k += 1
    ctx.sync()
    el = time.perf_counter() - t0

    # kernel durations: same number of steps again with event timing on
    ctx.profile_enable(True)
    for _ in range(args.steps):
        groom.step(wl.dt, tms[k], cols[k], prm[k])
        k += 1
    ctx.sync()
    n_launch, kern_ms = ctx.profile_read()
    ctx.profile_enable(False)
    out, _ = groom.read()
    if not np.isfinite(out).all():
        print("bench.py: non-finite state after the run", file=sys.stderr)
        sys.exit(1)
    groom.destroy()
    ctx.close()

    pts = g.S * g.N
    value = pts * args.steps / el
    avg_kernel_s = kern_ms / 1e3 / max(1, n_launch)
    bstep = scenes.algorithmic_bytes(g.S, g.N, g.grid_dim, wl.hair_hair)
    achieved = bstep / avg_kernel_s / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"traffic_{wl.name}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
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
                               f"wind + noise, orbiting sphere, friction {wl.friction}",
                   "strands": g.S, "points": g.N, "grid_dim": g.grid_dim, "hair_hair": wl.hair_hair},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel_avg_us": avg_kernel_s * 1e6, "algorithmic_bytes_per_launch": bstep},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
