"""The shader's commented-out wind forcing (compute.comp:150-151) on the GPU
against the oracle's FP32 restatement, bit for bit: the forcing's sin/cos/fbm
are one fixed binary32 algorithm on both sides (pinned against binary64 in
test_oracle.py)."""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from helpers import assert_bitwise, gpu_params, make_gpu_groom, oracle_params

pytestmark = pytest.mark.gpu


def _strip_corr(a, N):
    o = a.reshape(-1, 3, N, 4).copy()
    o[:, 2] = 0
    return o


@pytest.mark.parametrize("mode", [1, 2])
def test_forcing_bitwise(ctx, mode):
    wl = scenes.plain_100k(S=1500, N=16, G=32)
    g = wl.grooms[0]
    gpu = make_gpu_groom(ctx, wl)
    ref = g.strands.copy()
    grid = None
    try:
        for k in range(25):
            t = float(np.float32(wl.time(k) + 3.7 * k))  # totalTime sweeps many phases of the fbm
            cols = [scenes.sphere_collider(wl.sphere(k))]
            gpu.step(wl.dt, t, cols, gpu_params(wl, k, forcing=mode))
            grid = oracle.step32(oracle_params(wl, k, total_time=t, forcing=mode), ref, g.N, grid)
            if k % 5 == 4:
                out, _ = gpu.read()
                assert_bitwise(_strip_corr(out, g.N), _strip_corr(ref, g.N), f"mode {mode} step {k}", g.N)
        assert np.array_equal(gpu.read_grid(), grid.reshape(-1))
        # the forcing moved the strands: the same run without it differs
        plain = g.strands.copy()
        pg = None
        for k in range(25):
            pg = oracle.step32(oracle_params(wl, k), plain, g.N, pg)
        assert not np.array_equal(plain, ref)
    finally:
        gpu.destroy()


def test_forcing_in_a_batch_bitwise(ctx):
    """Two grooms in one batched launch, one per forcing mode."""
    wls = [scenes.plain_100k(S=700, N=12, G=16, seed=31), scenes.plain_100k(S=500, N=16, G=16, seed=32)]
    gpus = [make_gpu_groom(ctx, w) for w in wls]
    refs = [w.grooms[0].strands.copy() for w in wls]
    grids = [None, None]
    try:
        for k in range(10):
            items = []
            for j, w in enumerate(wls):
                t = float(np.float32(w.time(k) + 1.3 * k + j))
                cols = [scenes.sphere_collider(w.sphere(k))]
                items.append((gpus[j], w.dt, t, cols, gpu_params(w, k, forcing=j + 1)))
                grids[j] = oracle.step32(oracle_params(w, k, total_time=t, forcing=j + 1), refs[j],
                                         w.grooms[0].N, grids[j])
            ctx.step_batch(items)
        for j, w in enumerate(wls):
            out, _ = gpus[j].read()
            N = w.grooms[0].N
            assert_bitwise(_strip_corr(out, N), _strip_corr(refs[j], N), f"groom {j}", N)
    finally:
        for gp in gpus:
            gp.destroy()


def test_forcing_mode_out_of_range_is_rejected(ctx):
    wl = scenes.plain_1k(S=64)
    gpu = make_gpu_groom(ctx, wl)
    try:
        with pytest.raises(hairsim.HairSimError) as e:
            gpu.step(wl.dt, 0.0, [], hairsim.step_params(forcing=3))
        assert e.value.code == hairsim.HS_ERR_INVALID
    finally:
        gpu.destroy()


def test_forcing_with_mesh_and_ellipsoid_bitwise(ctx):
    """Forcing (mode 2), the mesh pass and an ellipsoid in one step: the
    force terms add in the shader's order (gravity + wind + noise, forcing,
    then the collider average with the mesh last)."""
    wl = scenes.mesh_head_scene("forcing-mesh", 41, 700, 12, 16, subdiv=3, margin=0.06, mesh_grid=(12, 14, 12),
                                steps=12)
    g = wl.grooms[0]
    V, T = wl.mesh
    mid = ctx.mesh(V, T, grid_dim=wl.mesh_grid, margin=wl.mesh_margin)
    recs = oracle.mesh_prepare(V, T)
    ell = hairsim.collider_from_trs((0.4, 0.9, 0.3), (0.0, 30.0, 10.0), (0.5, 0.4, 0.6)).as_array()
    gpu = make_gpu_groom(ctx, wl)
    ref = g.strands.copy()
    grid = None
    try:
        for k in range(12):
            t = float(np.float32(wl.time(k) + 2.1 * k))
            cols = [scenes.sphere_collider(wl.sphere(k)), ell]
            gpu.step(wl.dt, t, cols, gpu_params(wl, k, ncol=2, forcing=2, mesh_id=mid))
            grid = oracle.step32(oracle_params(wl, k, colliders=np.stack(cols), total_time=t, forcing=2, mesh=recs,
                                               mesh_margin=wl.mesh_margin), ref, g.N, grid)
        out, _ = gpu.read()
        assert_bitwise(_strip_corr(out, g.N), _strip_corr(ref, g.N), "forcing + mesh + ellipsoid", g.N)
    finally:
        gpu.destroy()
        ctx.mesh_destroy(mid)
