"""The batched and served workloads at their configured shapes, against ref32.

* batched-256: 256 grooms (S log-uniform in [512, 16384], N in {8,16,32,64},
  32^3 grids), stepped by captured-graph batched launches: every groom equals
  stepping it alone, four of them (the largest, an N = 64 one, the smallest,
  an N = 8 one) equal ref32 bit for bit, and no partial box fell back to
  device atomics or overflowed a pool.
* served: jobs with N = 64 and S up to 32768 on 32^3 grids, 60-90 steps each,
  keyframed wind and sphere, submitted together to a server with graphs on;
  every job's result equals ref32 stepping the same keyframed inputs.
"""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from helpers import assert_bitwise, gpu_params, make_gpu_groom, oracle_params

pytestmark = pytest.mark.gpu


def _no_corr(a, N):
    o = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3, N, 4).copy()
    o[:, 2] = 0
    return o


def test_batched_256_full_shape(ctx):
    wl = scenes.batched()
    n = len(wl.grooms)
    steps = 4
    print(f"batched-256: {n} grooms, {sum(g.S * g.N for g in wl.grooms)} points", flush=True)
    grs = [make_gpu_groom(ctx, wl, gi) for gi in range(n)]
    ctx.set_batch_graphs(True)
    try:
        b = ctx.batch(grs, max_colliders=1)
        b.params["damping"] = wl.damping
        b.params["sphere_radius"] = wl.sphere_radius
        b.params["friction"] = wl.friction
        b.params["hair_hair"] = 1
        b.params["wind"] = np.stack([wl.wind(0, gi) for gi in range(n)])
        b.set_colliders(np.stack([scenes.sphere_collider(wl.sphere(0, gi)) for gi in range(n)])[:, None])
        for k in range(steps):
            b.set_time(wl.dt, float(wl.time(k)))
            b.step()
        calls, caps = ctx.batch_stats()
        assert caps >= 2  # first step (no pending friction) and the steady shape
        for gi, gr in enumerate(grs):
            w = gr.grid_work()
            assert w["fallback"] == 0, (gi, w)
            assert w["pool_cells"] <= w["pool_capacity"] and w["boxes"] <= w["box_capacity"], (gi, w)
        print("batched-256: stepped, checking against ref32", flush=True)
        Ss = np.array([g.S for g in wl.grooms])
        Ns = np.array([g.N for g in wl.grooms])
        picks = {int(np.argmax(Ss)), int(np.argmin(Ss))}
        for N in (64, 8):
            cand = np.flatnonzero(Ns == N)
            picks.add(int(cand[np.argmax(Ss[cand])]))
        for gi in sorted(picks):
            g = wl.grooms[gi]
            ref = g.strands.copy()
            grid = None
            for k in range(steps):
                grid = oracle.step32_mt(oracle_params(wl, k, gi), ref, g.N, grid, 8)
            assert_bitwise(_no_corr(grs[gi].read()[0], g.N), _no_corr(ref, g.N), f"groom {gi} vs ref32", g.N)
            assert np.array_equal(grs[gi].read_grid(), grid.reshape(-1)), gi
            print(f"batched-256: groom {gi} (S {g.S}, N {g.N}) bitwise ref32", flush=True)
        # every groom equals stepping it alone
        for gi in range(n):
            g = wl.grooms[gi]
            al = make_gpu_groom(ctx, wl, gi)
            for k in range(steps):
                al.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k, gi))], gpu_params(wl, k, gi))
            assert_bitwise(grs[gi].read()[0], al.read()[0], f"groom {gi} batched vs alone", g.N)
            assert np.array_equal(grs[gi].read_grid(), al.read_grid()), gi
            al.destroy()
        print("batched-256: every groom equals stepping it alone", flush=True)
    finally:
        ctx.set_batch_graphs(False)
        for gr in grs:
            gr.destroy()


GRID32 = dict(grid_dim=32, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / 32)
DT = 1.0 / 60.0


def _served_jobs():
    rng = np.random.default_rng(scenes.SEED_BASE + 33)
    shapes = [(32768, 64, 60), (20000, 32, 75), (256, 64, 90), (12000, 64, 66), (5000, 16, 80), (900, 32, 61)]
    jobs = []
    for S, N, steps in shapes:
        roots, nrm = scenes.hemisphere_roots(rng, S)
        aos, seg = scenes.straight_strands(roots, nrm, N, rng.uniform(0.8, 1.2, S))
        nk = int(np.ceil(steps * DT / 0.25)) + 2
        keys = np.zeros((nk, 7), dtype=np.float32)
        keys[:, 0] = np.arange(nk) * 0.25
        keys[:, 1:4] = rng.normal(0.0, 1.5, (nk, 3))
        ang = rng.uniform(0, 2 * np.pi) + np.arange(nk) * 0.4
        keys[:, 4] = 1.1 * np.cos(ang)
        keys[:, 5] = rng.uniform(0.0, 1.0)
        keys[:, 6] = 1.1 * np.sin(ang)
        noise = (rng.standard_normal(S) * 0.1).astype(np.float32)
        jobs.append(dict(aos=aos, N=N, seg=seg, steps=steps, keys=keys, noise=noise))
    return jobs


def _ref32_job(job):
    N = job["N"]
    ref = np.ascontiguousarray(job["aos"], dtype=np.float32).reshape(-1).copy()
    karr = hairsim.keyframes(job["keys"])
    grid = None
    for k in range(job["steps"]):
        t = float(np.float32(k + 1) * np.float32(DT))
        wind, sph = hairsim.keyframes_eval(karr, t)
        p = oracle.Params(dt=DT, wind=wind, damping=0.9, sphere_radius=0.3, friction=0.05, hair_hair=True,
                          colliders=scenes.sphere_collider(sph), seg_length=job["seg"], wind_noise=job["noise"],
                          total_time=t, **GRID32)
        grid = oracle.step32_mt(p, ref, N, grid, 8)
    return ref


def test_served_full_shape_vs_ref32(ctx):
    jobs = _served_jobs()
    prm = hairsim.step_params(damping=0.9, sphere_radius=0.3, friction=0.05, num_colliders=1, hair_hair=True)
    print(f"served: {len(jobs)} jobs, {sum(j['aos'].size // 12 for j in jobs)} points", flush=True)
    with hairsim.Server(0, hbm_budget=8 << 30, max_points_per_step=12_000_000, use_graphs=True) as srv:
        ids = [srv.submit(j["aos"], j["N"], j["steps"], DT, j["keys"], prm, seg_length=j["seg"],
                          wind_noise=j["noise"], **GRID32)[0] for j in jobs]
        outs = [srv.wait(i, timeout=120.0) for i in ids]
        srv.drain()
        st = srv.stats()
        code, msg = srv.error()
    assert code == 0, msg
    assert st["completed"] == len(jobs)
    assert st["peak_jobs_per_launch"] >= 2
    print("served: done, checking against ref32", flush=True)
    for i, (j, out) in enumerate(zip(jobs, outs)):
        ref = _ref32_job(j)
        assert_bitwise(_no_corr(out, j["N"]), _no_corr(ref, j["N"]), f"job {i} (S {j['aos'].size // (12 * j['N'])}, "
                       f"N {j['N']}, {j['steps']} steps)", j["N"])
        print(f"served: job {i} bitwise ref32", flush=True)
