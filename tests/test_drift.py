"""FP32 against the FP64 oracle: per step on every collider and forcing path,
and free-running over a whole 240-step scene (the stated drift, tests/drift.py).

Per step (same FP32 state into both sides): positions within
1e-5 * max(|p|, 1), velocities within 1e-5 * max(|p|, 1) * damping / dt.
The GPU tests run the HIP library; the CPU tests run ref32, which the GPU
parity tests show to be bit-identical to it.
"""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from drift import check_drift, drift
from helpers import gpu_params, make_gpu_groom, oracle_params, positions, velocities

REF_DEMO = [((2.0, 0.0, 1.0), (0, 0, 0), (1, 1, 1)),
            ((0.0, 2.64, 0.08), (-38.270, 0, 0), (0.817, 1.158, 1.01)),
            ((0.0, 1.35, -0.288), (18.301, 0, 0), (0.457, 1.0, 0.538)),
            ((0.0, -0.380, -0.116), (-17.260, 0, 0), (1.078, 1.683, 0.974)),
            ((-0.698, 0.087, -0.36), (-20.254, 13.144, 34.5), (0.721, 1.0, 0.724)),
            ((0.698, 0.087, -0.36), (-20.254, 13.144, -34.5), (0.721, 1.0, 0.724))]


def _demo_colliders():
    """The six colliders of the reference demo (src/main.cpp:229-237), made by
    the library's host TRS helper."""
    return np.stack([hairsim.collider_from_trs(*s).as_array() for s in REF_DEMO])


def _path(name):
    """(workload, steps, per-step GPU/oracle extras) of one collider or forcing path."""
    if name == "ellipsoids":
        # reference constants, strands hanging through the five demo ellipsoids
        rng = np.random.default_rng(7)
        S, N = 900, 10
        roots, nrm = scenes.hemisphere_roots(rng, S)
        aos, seg = scenes.straight_strands(roots + np.array([0.0, 2.3, 0.0]), nrm, N, np.full(S, 2.5))
        g = scenes.Groom(aos, N, seg, None, grid_dim=64, grid_origin=(-3.0, -2.0, -5.0), grid_h=7.0 / 64.0)
        wl = scenes.Workload("reference-demo", dt=1.0 / 60.0, steps=20, damping=0.998, friction=0.08,
                             hair_hair=True, sphere_radius=1.0, grooms=[g])
        cols = _demo_colliders()
        return wl, 20, lambda k: dict(cols=cols, ncol=6, t=float(wl.time(k)), forcing=0)
    if name in ("forcing1", "forcing2"):
        wl = scenes.plain_100k(S=1500, N=16, G=32)
        mode = int(name[-1])
        return wl, 25, lambda k: dict(cols=scenes.sphere_collider(wl.sphere(k))[None], ncol=1,
                                      t=float(np.float32(wl.time(k) + 3.7 * k)), forcing=mode)
    if name == "mesh":
        wl = scenes.mesh_head_scene("mesh-drift", 3, 1024, 16, 32, subdiv=3, margin=0.06, mesh_grid=(12, 14, 12),
                                    steps=30)
        return wl, 30, lambda k: dict(cols=scenes.sphere_collider(wl.sphere(k))[None], ncol=1,
                                      t=float(wl.time(k)), forcing=0)
    raise ValueError(name)


PATHS = ["ellipsoids", "forcing1", "forcing2", "mesh"]


def _oracle_kw(wl, x, recs):
    kw = dict(colliders=x["cols"], total_time=x["t"], forcing=x["forcing"])
    if recs is not None:
        kw.update(mesh=recs, mesh_margin=wl.mesh_margin)
    return kw


def _compare(wl, before, after, k, recs, x):
    """oracle64 one step from `before` (FP32 state), against the FP32 `after`."""
    g = wl.grooms[0]
    s64 = before.astype(np.float64)
    oracle.step64(oracle_params(wl, k, **_oracle_kw(wl, x, recs)), s64, g.N)
    p64, p32 = positions(s64, g.N), positions(after, g.N)
    v64, v32 = velocities(s64, g.N), velocities(after, g.N)
    scale = np.maximum(np.abs(p64), 1.0)
    return (float(np.max(np.abs(p32 - p64) / scale)),
            float(np.max(np.abs(v32 - v64) / (scale * wl.damping / wl.dt))))


@pytest.mark.parametrize("name", PATHS)
def test_ref32_vs_oracle64_per_step_paths(name):
    wl, steps, extra = _path(name)
    g = wl.grooms[0]
    recs = None if wl.mesh is None else oracle.mesh_prepare(*wl.mesh)
    s32 = g.strands.copy()
    worst_p = worst_v = 0.0
    for k in range(steps):
        x = extra(k)
        before = s32.copy()
        oracle.step32_mt(oracle_params(wl, k, **_oracle_kw(wl, x, recs)), s32, g.N, None, 8)
        ep, ev = _compare(wl, before, s32, k, recs, x)
        worst_p, worst_v = max(worst_p, ep), max(worst_v, ev)
    assert worst_p <= 1e-5, worst_p
    assert worst_v <= 1e-5, worst_v


@pytest.mark.gpu
@pytest.mark.parametrize("name", PATHS)
def test_gpu_vs_oracle64_per_step_paths(ctx, name):
    wl, steps, extra = _path(name)
    recs = None
    mid = 0
    if wl.mesh is not None:
        recs = oracle.mesh_prepare(*wl.mesh)
        mid = ctx.mesh(*wl.mesh, grid_dim=wl.mesh_grid, margin=wl.mesh_margin)
    gpu = make_gpu_groom(ctx, wl)
    worst_p = worst_v = 0.0
    try:
        for k in range(steps):
            x = extra(k)
            before, _ = gpu.read()
            gpu.step(wl.dt, x["t"], list(x["cols"]),
                     gpu_params(wl, k, ncol=x["ncol"], forcing=x["forcing"], mesh_id=mid))
            after, _ = gpu.read()
            ep, ev = _compare(wl, before, after, k, recs, x)
            worst_p, worst_v = max(worst_p, ep), max(worst_v, ev)
    finally:
        gpu.destroy()
        if mid:
            ctx.mesh_destroy(mid)
    assert worst_p <= 1e-5, worst_p
    assert worst_v <= 1e-5, worst_v


# ------------------------------------------------ free-running, 240 steps --
def _ref32_stepper(wl):
    g = wl.grooms[0]
    s32 = g.strands.copy()
    st = {"grid": None}

    def step(k, p):
        st["grid"] = oracle.step32(p, s32, g.N, st["grid"])
    return step, lambda: s32


@pytest.mark.parametrize("which", ["plain-1k", "plain-100k-1024"])
def test_ref32_drift_240_steps(which):
    wl = scenes.plain_1k() if which == "plain-1k" else scenes.plain_100k(S=1024, N=32, G=64)
    step, read = _ref32_stepper(wl)
    check_drift(*drift(wl, 240, step, read))


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["plain-1k", "plain-100k-2048"])
def test_gpu_drift_240_steps(ctx, which):
    """The GPU and oracle64 run free from the same seeded state for the whole
    240-step scene (the 1024 x 16 scene in full; the 100k x 32 scene's
    physics on 2048 strands with its 64^3 grid)."""
    wl = scenes.plain_1k() if which == "plain-1k" else scenes.plain_100k(S=2048, N=32, G=64)
    gpu = make_gpu_groom(ctx, wl)
    state = {}

    def step(k, p):
        gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
        state.pop("aos", None)

    def read():
        if "aos" not in state:
            state["aos"] = gpu.read()[0]
        return state["aos"]
    try:
        check_drift(*drift(wl, 240, step, read))
    finally:
        gpu.destroy()
