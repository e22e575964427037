"""Batched stepping (hs_step_batch): many grooms of different shapes in one
step launch + one reduce launch give each groom exactly the result of
stepping it alone, and the oracle's."""
import numpy as np
import pytest

import hairsim
from hairsim import scenes
from helpers import assert_bitwise, gpu_params, make_gpu_groom, run_parity

pytestmark = pytest.mark.gpu


def small_batch(n=12, seed=5):
    return scenes.batched(seed=seed, n_grooms=n, s_lo=64, s_hi=3000, G=16)


def test_batch_equals_alone(ctx):
    wl = small_batch()
    steps = 6
    alone = []
    for gi in range(len(wl.grooms)):
        gr = make_gpu_groom(ctx, wl, gi)
        for k in range(steps):
            gr.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k, gi))], gpu_params(wl, k, gi))
        alone.append((gr.read()[0], gr.read_grid()))
        gr.destroy()
    grs = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    for k in range(steps):
        ctx.step_batch([(gr, wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k, gi))],
                         gpu_params(wl, k, gi)) for gi, gr in enumerate(grs)])
    for gi, gr in enumerate(grs):
        out, _ = gr.read()
        assert_bitwise(out, alone[gi][0], f"groom {gi}", wl.grooms[gi].N)
        assert np.array_equal(gr.read_grid(), alone[gi][1]), gi
        gr.destroy()


def test_batch_matches_oracle_for_one_groom(ctx):
    """The batched path against ref32 directly (first groom of a batch, the rest as load)."""
    wl = small_batch(n=5, seed=9)
    ref, _ = run_parity(ctx, wl, 4, groom=0)  # alone, vs oracle
    grs = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    for k in range(4):
        ctx.step_batch([(gr, wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k, gi))],
                         gpu_params(wl, k, gi)) for gi, gr in enumerate(grs)])
    out, _ = grs[0].read()
    o = out.reshape(-1, 3, wl.grooms[0].N, 4).copy()
    r = ref.reshape(-1, 3, wl.grooms[0].N, 4).copy()
    o[:, 2] = 0
    r[:, 2] = 0
    assert_bitwise(o, r, "batched groom 0 vs ref32", wl.grooms[0].N)
    for gr in grs:
        gr.destroy()


def test_batch_rejects_duplicates(ctx):
    wl = small_batch(n=2)
    gr = make_gpu_groom(ctx, wl, 0)
    p = gpu_params(wl, 0, 0)
    with pytest.raises(hairsim.HairSimError) as e:
        ctx.step_batch([(gr, wl.dt, 0.0, [scenes.sphere_collider(wl.sphere(0, 0))], p)] * 2)
    assert e.value.code == hairsim.HS_ERR_INVALID
    gr.destroy()
