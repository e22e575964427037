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


def _run_batch_object(ctx, wl, grs, steps, graphs, reads_between=False):
    ctx.set_batch_graphs(graphs)
    b = ctx.batch(grs, max_colliders=1)
    snaps = []
    for k in range(steps):
        b.set_time(wl.dt, float(wl.time(k)))
        b.set_colliders(np.stack([scenes.sphere_collider(wl.sphere(k, gi)) for gi in range(len(grs))])[:, None])
        b.params["wind"] = np.stack([wl.wind(k, gi) for gi in range(len(grs))])
        b.params["damping"] = wl.damping
        b.params["sphere_radius"] = wl.sphere_radius
        b.params["friction"] = wl.friction
        b.params["hair_hair"] = 1
        b.step()
        if reads_between:
            # eager work and a synchronisation between graph replays
            snaps.append(grs[k % len(grs)].read()[0])
    ctx.set_batch_graphs(False)
    return snaps


def test_batch_graph_replays_equal_eager(ctx):
    """Captured-graph batch steps: every replay gives the eager result bit for bit.

    Replays alternate between the two pinned argument slots, with reads (device
    to host copies + synchronisation) between them.
    """
    wl = small_batch(n=9, seed=21)
    steps = 7
    eager = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    snaps_e = _run_batch_object(ctx, wl, eager, steps, graphs=False, reads_between=True)
    calls0, caps0 = ctx.batch_stats()
    graph = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    snaps_g = _run_batch_object(ctx, wl, graph, steps, graphs=True, reads_between=True)
    calls1, caps1 = ctx.batch_stats()
    assert calls1 - calls0 == steps
    # shapes: first step (no pending friction) and the rest; one graph per pinned slot
    assert caps1 - caps0 == 3
    for k, (a, b) in enumerate(zip(snaps_e, snaps_g)):
        assert_bitwise(b, a, f"replay {k}", wl.grooms[k % len(wl.grooms)].N)
    for gi in range(len(wl.grooms)):
        assert_bitwise(graph[gi].read()[0], eager[gi].read()[0], f"groom {gi}", wl.grooms[gi].N)
        assert np.array_equal(graph[gi].read_grid(), eager[gi].read_grid()), gi
    # the batch object path equals the list path / stepping alone
    for gi in range(len(wl.grooms)):
        gr = make_gpu_groom(ctx, wl, gi)
        for k in range(steps):
            gr.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k, gi))], gpu_params(wl, k, gi))
        assert_bitwise(eager[gi].read()[0], gr.read()[0], f"alone {gi}", wl.grooms[gi].N)
        gr.destroy()
    for gr in eager + graph:
        gr.destroy()


def test_batch_ellipsoids_in_blob(ctx):
    """Grooms with ellipsoid colliders in a batch read their own colliders from the argument blob."""
    wl = small_batch(n=4, seed=33)
    demo = [scenes.sphere_collider((0.0, 0.0, 1.2))]
    for j, t in enumerate([(0.3, 0.8, 0.0), (-0.5, 0.2, 0.4), (0.0, -0.3, -0.6)]):
        demo.append(hairsim.collider_from_trs(t, (10.0 * j, 20.0, 5.0), (0.6, 0.9, 0.7)).as_array())
    def cols_for(gi, k):
        return [scenes.sphere_collider(wl.sphere(k, gi))] + demo[1:1 + (gi % 4)]
    alone = []
    for gi in range(len(wl.grooms)):
        gr = make_gpu_groom(ctx, wl, gi)
        for k in range(4):
            cl = cols_for(gi, k)
            gr.step(wl.dt, float(wl.time(k)), cl, gpu_params(wl, k, gi, ncol=len(cl)))
        alone.append(gr.read()[0])
        gr.destroy()
    for graphs in (False, True):
        ctx.set_batch_graphs(graphs)
        grs = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
        for k in range(4):
            ctx.step_batch([(gr, wl.dt, float(wl.time(k)), cols_for(gi, k),
                             gpu_params(wl, k, gi, ncol=len(cols_for(gi, k)))) for gi, gr in enumerate(grs)])
        for gi, gr in enumerate(grs):
            assert_bitwise(gr.read()[0], alone[gi], f"groom {gi} graphs={graphs}", wl.grooms[gi].N)
            gr.destroy()
    ctx.set_batch_graphs(False)


def test_autotune_leaves_state_and_results_unchanged(ctx):
    """hs_autotune_batch runs the candidates on copies: the grooms' state is
    untouched, and stepping with any tile gives the same bits."""
    wl = small_batch(n=6, seed=41)
    grs = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    ref = [make_gpu_groom(ctx, wl, gi) for gi in range(len(wl.grooms))]
    b = ctx.batch(grs)
    b.params["damping"] = wl.damping
    b.params["sphere_radius"] = wl.sphere_radius
    b.params["friction"] = wl.friction
    b.params["hair_hair"] = 1
    b.set_colliders(np.stack([scenes.sphere_collider(wl.sphere(0, gi)) for gi in range(len(grs))])[:, None])
    b.set_time(wl.dt, float(wl.time(0)))
    b.step()  # a pending friction and a non-trivial parity before tuning
    before = [g.read()[0] for g in grs]
    tile, ms = ctx.autotune(b, steps=2)
    assert ms > 0.0 and ctx.get_tile() == tile
    for gi, g in enumerate(grs):
        assert_bitwise(g.read()[0], before[gi], f"groom {gi} after autotune", wl.grooms[gi].N)
    for k in range(1, 4):
        b.set_time(wl.dt, float(wl.time(k)))
        b.step()
    ctx.set_tile(2048, 2048, 1)
    rb = ctx.batch(ref)
    rb.params[:] = b.params
    rb.colliders[:] = b.colliders
    rb.items["num_colliders"] = b.items["num_colliders"]
    for k in range(0, 4):
        rb.set_time(wl.dt, float(wl.time(k)))
        rb.step()
    for gi in range(len(grs)):
        assert_bitwise(grs[gi].read()[0], ref[gi].read()[0], f"tuned vs default groom {gi}", wl.grooms[gi].N)
        assert np.array_equal(grs[gi].read_grid(), ref[gi].read_grid())
    for g in grs + ref:
        g.destroy()


@pytest.mark.parametrize("tile", [(512, 512, 0), (1024, 2048, 1), (2048, 512, 0), (64, 64, 1),
                                  (2048, 2048, 1, 256, 8), (1024, 512, 0, 256, 12), (64, 64, 1, 256, 8),
                                  (2048, 1024, 1, 256, 12)])
def test_tile_choices_bitwise(ctx, tile):
    wl = scenes.plain_100k(S=3000, N=20, G=32)
    ctx.set_tile(*tile)
    try:
        assert ctx.get_tile()[:len(tile)] == tile
        run_parity(ctx, wl, 5)
    finally:
        ctx.set_tile(2048, 2048, 1)


def test_tile_shape_changes_bitwise(ctx):
    """The tile shape may change between steps: the step after a change
    gathers from the float grid (the last step's boxes are laid out for the
    old shape) and the deferred-list halves keep one stride, so strands and
    voxel ints stay equal to ref32.  3000 strands (not a multiple of 256) x 20
    points: 24 or 12 workgroups, 2 or 3 segments."""
    wl = scenes.plain_100k(S=3000, N=20, G=32)
    shapes = hairsim.TILE_SHAPES
    order = [0, 1, 2, 0, 2, 1, 1, 0]
    try:
        run_parity(ctx, wl, len(order), tile_fn=lambda k: (2048, 1024 if k % 2 else 2048, k % 2) + shapes[order[k]])
        # and in a batch, with small splat windows (many deferred segments)
        bw = small_batch(n=5, seed=61)
        ctx.set_tile(2048, 2048, 1)
        alone = [run_parity(ctx, bw, 6, groom=gi)[0] for gi in range(len(bw.grooms))]
        grs = [make_gpu_groom(ctx, bw, gi) for gi in range(len(bw.grooms))]
        for k in range(6):
            ctx.set_tile(2048, 256 if k % 2 else 2048, 1, *shapes[order[k]])
            ctx.step_batch([(gr, bw.dt, float(bw.time(k)), [scenes.sphere_collider(bw.sphere(k, gi))],
                             gpu_params(bw, k, gi)) for gi, gr in enumerate(grs)])
        for gi, gr in enumerate(grs):
            o = gr.read()[0].reshape(-1, 3, bw.grooms[gi].N, 4).copy()
            r = alone[gi].reshape(-1, 3, bw.grooms[gi].N, 4).copy()
            o[:, 2] = 0
            r[:, 2] = 0
            assert_bitwise(o, r, f"batched groom {gi} with shape changes vs ref32", bw.grooms[gi].N)
            gr.destroy()
    finally:
        ctx.set_tile(2048, 2048, 1)


def test_set_tile_rejects_unbuilt_shape(ctx):
    for bad in [(2048, 2048, 1, 64, 16), (2048, 2048, 1, 128, 8), (2048, 2048, 1, 256, 16), (2048, 2048, 1, 0, 0)]:
        with pytest.raises(hairsim.HairSimError) as e:
            ctx.set_tile(*bad)
        assert e.value.code == hairsim.HS_ERR_INVALID
    assert ctx.get_tile() == (2048, 2048, 1, 128, 16)


def test_deferred_chunks_bitwise(ctx):
    """Boxes larger than one splat window are splatted chunk by chunk by the
    workgroups that finish first (and by their owners): with 256-cell windows
    many workgroup-segments are deferred, in a single groom and in a batch,
    and the strands and voxel ints stay equal to ref32."""
    wl = scenes.plain_100k(S=5000, N=32, G=64)
    ctx.set_tile(2048, 256, 1)
    try:
        gpu = make_gpu_groom(ctx, wl)
        for k in range(2):
            gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
        w = gpu.grid_work()
        gpu.destroy()
        # (with windows this small the shared pool can fill, and the rest goes
        # to the overflow grid: also exact)
        assert w["deferred"] > 0, w
        run_parity(ctx, wl, 6, check_every=2)
        bw = small_batch(n=6, seed=51)
        alone = []
        for gi in range(len(bw.grooms)):
            alone.append(run_parity(ctx, bw, 3, groom=gi)[0])
        grs = [make_gpu_groom(ctx, bw, gi) for gi in range(len(bw.grooms))]
        for k in range(3):
            ctx.step_batch([(gr, bw.dt, float(bw.time(k)), [scenes.sphere_collider(bw.sphere(k, gi))],
                             gpu_params(bw, k, gi)) for gi, gr in enumerate(grs)])
        for gi, gr in enumerate(grs):
            o = gr.read()[0].reshape(-1, 3, bw.grooms[gi].N, 4).copy()
            r = alone[gi].reshape(-1, 3, bw.grooms[gi].N, 4).copy()
            o[:, 2] = 0
            r[:, 2] = 0
            assert_bitwise(o, r, f"batched groom {gi} vs ref32", bw.grooms[gi].N)
            gr.destroy()
    finally:
        ctx.set_tile(2048, 2048, 1)
