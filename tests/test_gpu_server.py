"""Served jobs (hs_server_*): continuous batching over the batched step.

Jobs of different shapes and lengths arrive while others run; the admission
budget forces some to wait.  Each job's final state must equal stepping that
job alone with hs_step on the same keyframed inputs, bit for bit (the batched
kernels compute every groom exactly as the single-groom kernels do).
"""
import time

import numpy as np
import pytest

import hairsim
from hairsim import scenes
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _job_set(seed, n):
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n):
        S = int(rng.integers(200, 2500))
        N = int(rng.choice([8, 16, 32]))
        roots, nrm = scenes.hemisphere_roots(rng, S)
        L = rng.uniform(0.8, 1.2, S)
        aos, seg = scenes.straight_strands(roots, nrm, N, L)
        steps = int(rng.integers(4, 14))
        keys = np.zeros((3, 7), dtype=np.float32)
        keys[:, 0] = [0.0, 0.1, 0.25]
        keys[:, 1:4] = rng.normal(0.0, 1.5, (3, 3))
        keys[:, 4:7] = rng.uniform(-1.2, 1.2, (3, 3))
        noise = rng.normal(0.0, 0.1, S).astype(np.float32) if j % 2 else None
        jobs.append(dict(aos=aos, N=N, seg=seg, steps=steps, keys=keys, noise=noise))
    return jobs


def _params(ncol=1, forcing=0):
    return hairsim.step_params(damping=0.9, sphere_radius=0.3, friction=0.05, num_colliders=ncol, hair_hair=True,
                               forcing=forcing)


GRID = dict(grid_dim=16, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / 16)
DT = 1.0 / 60.0


def _alone(ctx, job, forcing=0):
    gr = ctx.groom(job["aos"], job["N"], seg_length=job["seg"], wind_noise=job["noise"], **GRID)
    karr = hairsim.keyframes(job["keys"])
    for k in range(job["steps"]):
        t = float(np.float32(k + 1) * np.float32(DT))
        wind, sph = hairsim.keyframes_eval(karr, t)
        p = hairsim.step_params(wind=wind, damping=0.9, sphere_radius=0.3, friction=0.05, num_colliders=1,
                                hair_hair=True, forcing=forcing)
        gr.step(DT, t, [scenes.sphere_collider(sph)], p)
    out, _ = gr.read()
    gr.destroy()
    return out


def test_served_jobs_equal_alone(ctx):
    jobs = _job_set(77, 10)
    # a third of the jobs each with no forcing, the sinusoidal and the fbm wind
    refs = [_alone(ctx, j, forcing=i % 3) for i, j in enumerate(jobs)]
    sizes = [hairsim.groom_bytes(j["aos"].size // (12 * j["N"]), j["N"], 16) + j["aos"].nbytes for j in jobs]
    budget = int(sum(sorted(sizes)[-3:]) * 1.05)  # about three jobs resident at once
    # one step loop: with several, each of the few resident jobs would have a loop to itself
    with hairsim.Server(0, hbm_budget=budget, max_points_per_step=60000, step_loops=1, launch_depth=2) as srv:
        ids = []
        for i, j in enumerate(jobs):
            jid, _ = srv.submit(j["aos"], j["N"], j["steps"], DT, j["keys"], _params(forcing=i % 3),
                                seg_length=j["seg"], wind_noise=j["noise"], **GRID)
            ids.append(jid)
            if i % 3 == 2:
                time.sleep(0.01)  # later jobs arrive while earlier ones run
        for i, jid in enumerate(ids):
            out = srv.wait(jid, timeout=60.0)
            assert_bitwise(out, refs[i], f"job {i}", jobs[i]["N"])
        srv.drain()
        st = srv.stats()
    assert st["completed"] == len(jobs)
    assert st["peak_resident_bytes"] <= budget
    assert st["launches"] >= max(j["steps"] for j in jobs)
    # more than one job per launch happened (batching), and not all at once (budget)
    assert 1 < st["peak_jobs_per_launch"] < len(jobs)
    assert st["job_steps"] == sum(j["steps"] for j in jobs)
    assert st["p99_step_ms"] >= st["p50_step_ms"] > 0.0
    assert st["p99_latency_ms"] >= st["p50_latency_ms"] >= st["p50_step_ms"] * 0.5
    assert 0.0 < st["gpu_busy_s"] <= st["busy_s"] * 1.05


@pytest.mark.parametrize("loops,depth", [(1, 2), (3, 1), (2, 2)])
def test_round_robin_launches_equal_alone(ctx, loops, depth):
    """Every job resident at once (a large HBM budget) but a point limit of
    about three jobs per launch: launches take jobs in round-robin order, up
    to `depth` launches are queued per step loop, `loops` loops launch
    concurrently on their own streams, and every job still ends bit-identical
    to stepping it alone."""
    jobs = _job_set(91, 8)
    refs = [_alone(ctx, j, forcing=i % 3) for i, j in enumerate(jobs)]
    pts = [j["aos"].size // 12 for j in jobs]
    cap = int(sum(sorted(pts)[-3:]))
    with hairsim.Server(0, hbm_budget=1 << 30, max_points_per_step=cap, step_loops=loops, launch_depth=depth) as srv:
        ids = [srv.submit(j["aos"], j["N"], j["steps"], DT, j["keys"], _params(forcing=i % 3), seg_length=j["seg"],
                          wind_noise=j["noise"], **GRID)[0] for i, j in enumerate(jobs)]
        for i, jid in enumerate(ids):
            assert_bitwise(srv.wait(jid, timeout=60.0), refs[i], f"job {i}", jobs[i]["N"])
        srv.drain()
        st = srv.stats()
    assert st["completed"] == len(jobs)
    assert st["job_steps"] == sum(j["steps"] for j in jobs)
    assert st["point_updates"] == sum(p * j["steps"] for p, j in zip(pts, jobs))
    # launches never exceeded the point limit, so they took turns
    assert st["point_updates"] <= st["launches"] * cap
    assert st["launches"] > max(j["steps"] for j in jobs)


def test_server_rejects_and_cancels(ctx):
    jobs = _job_set(5, 4)
    j = jobs[0]
    small = hairsim.groom_bytes(j["aos"].size // (12 * j["N"]), j["N"], 16)
    with hairsim.Server(0, hbm_budget=small // 2, max_points_per_step=1 << 24) as srv:
        with pytest.raises(hairsim.HairSimError) as e:
            srv.submit(j["aos"], j["N"], 3, DT, j["keys"], _params(), seg_length=j["seg"], **GRID)
        assert e.value.code == hairsim.HS_ERR_BUDGET
    with hairsim.Server(0, hbm_budget=1 << 30, max_points_per_step=1) as srv:
        with pytest.raises(hairsim.HairSimError) as e:
            srv.submit(j["aos"], j["N"], 3, DT, j["keys"], _params(), seg_length=j["seg"], **GRID)
        assert e.value.code == hairsim.HS_ERR_BUDGET
    # a one-job-at-a-time server (the budget holds one job): queued jobs can
    # be cancelled; the admitted one finishes
    one = max(hairsim.groom_bytes(x["aos"].size // (12 * x["N"]), x["N"], 16) + 2 * x["aos"].nbytes + 512
              for x in jobs)
    with hairsim.Server(0, hbm_budget=one, max_points_per_step=1 << 24) as srv:
        ids = [srv.submit(x["aos"], x["N"], 40, DT, x["keys"], _params(), seg_length=x["seg"], **GRID)[0]
               for x in jobs]
        n = srv.cancel_pending()
        srv.drain()
        st = srv.stats()
        assert st["completed"] + n == len(jobs)
        assert st["cancelled"] == n
        done = cancelled = 0
        for i in ids:
            try:
                srv.wait(i, 0.0)
                done += 1
            except hairsim.HairSimError as e:
                assert e.code == hairsim.HS_ERR_NOT_FOUND
                cancelled += 1
        assert (done, cancelled) == (st["completed"], n)


def test_bad_job_rejected_at_submit_others_complete(ctx):
    """A job hs_step_batch would refuse (forcing 3, a dt that is not finite and
    positive) is refused at submit; the jobs submitted around it run to
    completion bit-identical to stepping them alone."""
    jobs = _job_set(31, 3)
    refs = [_alone(ctx, j) for j in jobs]
    with hairsim.Server(0, hbm_budget=1 << 30, max_points_per_step=1 << 22) as srv:
        ids = [srv.submit(jobs[0]["aos"], jobs[0]["N"], jobs[0]["steps"], DT, jobs[0]["keys"], _params(),
                          seg_length=jobs[0]["seg"], wind_noise=jobs[0]["noise"], **GRID)[0]]
        bad = jobs[1]
        for dt, forcing in ((DT, 3), (float("inf"), 0), (float("nan"), 0), (-DT, 0)):
            with pytest.raises(hairsim.HairSimError) as e:
                srv.submit(bad["aos"], bad["N"], 5, dt, bad["keys"], _params(forcing=forcing),
                           seg_length=bad["seg"], **GRID)
            assert e.value.code == hairsim.HS_ERR_INVALID
        for j in jobs[1:]:
            ids.append(srv.submit(j["aos"], j["N"], j["steps"], DT, j["keys"], _params(), seg_length=j["seg"],
                                  wind_noise=j["noise"], **GRID)[0])
        for i, jid in enumerate(ids):
            assert_bitwise(srv.wait(jid, timeout=60.0), refs[i], f"job {i}", jobs[i]["N"])
        srv.drain()
        code, _ = srv.error()
        assert code == 0
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["submitted"] == len(jobs)


def test_server_tile_and_loops_config(ctx):
    """An explicit tile for the step loops (here the library's 128 x 16) gives
    the same bits as the server's default tile and as stepping alone; a tile
    shape that is not built, a launch depth above 2 or too many loops are
    refused at creation."""
    jobs = _job_set(13, 3)
    refs = [_alone(ctx, j) for j in jobs]
    for kw in (dict(tile=(2048, 1024, 1, 128, 16), step_loops=2, launch_depth=2), dict(step_loops=1, launch_depth=1)):
        with hairsim.Server(0, hbm_budget=1 << 30, max_points_per_step=1 << 22, **kw) as srv:
            ids = [srv.submit(j["aos"], j["N"], j["steps"], DT, j["keys"], _params(), seg_length=j["seg"],
                              wind_noise=j["noise"], **GRID)[0] for j in jobs]
            for i, jid in enumerate(ids):
                assert_bitwise(srv.wait(jid, timeout=60.0), refs[i], f"job {i} {kw}", jobs[i]["N"])
    for bad in (dict(tile=(2048, 2048, 1, 64, 16)), dict(launch_depth=3), dict(step_loops=9)):
        with pytest.raises(hairsim.HairSimError) as e:
            hairsim.Server(0, hbm_budget=1 << 30, max_points_per_step=1 << 22, **bad)
        assert e.value.code == hairsim.HS_ERR_INVALID, bad
