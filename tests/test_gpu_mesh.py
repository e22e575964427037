"""Mesh collider on the GPU against the oracle's brute-force restatement.

The GPU looks up candidate triangles in a uniform grid; the oracle scans every
triangle.  Both apply the same contact rule (nearest point within the margin,
inside by the pseudo-normal test) in the same binary32 operations, so the
strands must agree bit for bit.
"""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _scene(S=1024, N=16, G=32, subdiv=3, seed=3):
    return scenes.mesh_head_scene("mesh-test", seed, S, N, G, subdiv=subdiv, margin=0.06, mesh_grid=(12, 14, 12),
                                  steps=40)


def _run(ctx, wl, steps, extra_cols=(), batch_with=None):
    g = wl.grooms[0]
    V, T = wl.mesh
    mid = ctx.mesh(V, T, grid_dim=wl.mesh_grid, margin=wl.mesh_margin)
    recs = oracle.mesh_prepare(V, T)
    gpu = ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, grid_dim=g.grid_dim,
                    grid_origin=g.grid_origin, grid_h=g.grid_h)
    other = None
    if batch_with is not None:
        og = batch_with.grooms[0]
        other = ctx.groom(og.strands, og.N, seg_length=og.seg, grid_dim=og.grid_dim, grid_origin=og.grid_origin,
                          grid_h=og.grid_h)
    ref = g.strands.copy()
    grid = None
    contacts = 0
    try:
        for k in range(steps):
            cols = [scenes.sphere_collider(wl.sphere(k))] + list(extra_cols)
            p = hairsim.step_params(wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                                    friction=wl.friction, num_colliders=len(cols), hair_hair=True, mesh_id=mid)
            if other is None:
                gpu.step(wl.dt, float(wl.time(k)), cols, p)
            else:
                po = hairsim.step_params(damping=0.9, num_colliders=1, sphere_radius=0.3, hair_hair=True)
                ctx.step_batch([(gpu, wl.dt, float(wl.time(k)), cols, p),
                                (other, wl.dt, float(wl.time(k)), [scenes.sphere_collider((0, 0, 1.2))], po)])
            op = oracle.Params(dt=wl.dt, wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                               friction=wl.friction, hair_hair=True, grid_dim=g.grid_dim, grid_origin=g.grid_origin,
                               grid_h=g.grid_h, colliders=np.stack(cols), seg_length=g.seg, wind_noise=g.noise,
                               noise_dir=g.noise_dir, mesh=recs, mesh_margin=wl.mesh_margin)
            if k % 10 == 9:
                # the mesh is in play: some points are in contact
                P = ref.reshape(-1, 3, g.N, 4)[:, 0, 1:, :3].reshape(-1, 3)
                hit = 0
                for x in P[:: max(1, len(P) // 400)]:
                    hit += oracle.force32(op, x)[0] > (1 if np.linalg.norm(x - cols[0][12:15]) < 0.3 else 0)
                contacts += hit
            grid = oracle.step32_mt(op, ref, g.N, grid, 8)
        out, _ = gpu.read()
        o = out.reshape(-1, 3, g.N, 4).copy()
        r = ref.reshape(-1, 3, g.N, 4).copy()
        o[:, 2] = 0
        r[:, 2] = 0
        assert_bitwise(o, r, "strands", g.N)
        assert np.array_equal(gpu.read_grid(), grid.reshape(-1))
    finally:
        gpu.destroy()
        if other is not None:
            other.destroy()
        ctx.mesh_destroy(mid)
    return contacts


def test_mesh_collider_bitwise(ctx):
    wl = _scene()
    contacts = _run(ctx, wl, 40)
    assert contacts > 0


def test_mesh_with_ellipsoids_bitwise(ctx):
    wl = _scene(S=700, N=12, seed=8)
    ell = hairsim.collider_from_trs((0.4, 0.9, 0.3), (0.0, 30.0, 10.0), (0.5, 0.4, 0.6)).as_array()
    _run(ctx, wl, 20, extra_cols=[ell])


def test_mesh_groom_in_batch_bitwise(ctx):
    wl = _scene(S=900, N=16, seed=12)
    _run(ctx, wl, 15, batch_with=scenes.plain_100k(S=600, N=8, G=16))


def test_mesh_errors(ctx):
    V, T = scenes.icosphere(1)
    with pytest.raises(hairsim.HairSimError) as e:
        ctx.mesh(V, T[:-1])
    assert e.value.code == hairsim.HS_ERR_INVALID
    Vn = V.copy()
    Vn[3, 1] = np.nan
    with pytest.raises(hairsim.HairSimError) as e:
        ctx.mesh(Vn, T)
    assert e.value.code == hairsim.HS_ERR_INVALID
    wl = scenes.plain_1k(S=64)
    g = wl.grooms[0]
    gpu = ctx.groom(g.strands, g.N)
    with pytest.raises(hairsim.HairSimError) as e:
        gpu.step(wl.dt, 0.0, [], hairsim.step_params(mesh_id=999))
    assert e.value.code == hairsim.HS_ERR_NOT_FOUND
    gpu.destroy()


@pytest.mark.parametrize("graphs", [False, True])
def test_two_mesh_grooms_batched_bitwise(ctx, graphs):
    """Two grooms of different shapes share one mesh in one batch; the mesh
    pass covers both.  Under graph replay the spring constant changes every
    step: the captured mesh pass must take each call's value."""
    wls = [_scene(S=600, N=12, seed=21), _scene(S=500, N=16, seed=22)]
    V, T = wls[0].mesh
    mid = ctx.mesh(V, T, grid_dim=wls[0].mesh_grid, margin=wls[0].mesh_margin)
    recs = oracle.mesh_prepare(V, T)
    gs = [w.grooms[0] for w in wls]
    gpus = [ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, grid_dim=g.grid_dim,
                      grid_origin=g.grid_origin, grid_h=g.grid_h) for g in gs]
    refs = [g.strands.copy() for g in gs]
    grids = [None, None]
    ctx.set_batch_graphs(graphs)
    try:
        for k in range(12):
            items, ops = [], []
            for j, (w, g) in enumerate(zip(wls, gs)):
                cols = [scenes.sphere_collider(w.sphere(k))]
                kk = 1900.0 + 150.0 * ((k + j) % 4)
                items.append((gpus[j], w.dt, float(w.time(k)), cols,
                              hairsim.step_params(wind=w.wind(k), damping=w.damping, spring_k=kk,
                                                  sphere_radius=w.sphere_radius, friction=w.friction,
                                                  num_colliders=1, hair_hair=True, mesh_id=mid)))
                ops.append(oracle.Params(dt=w.dt, wind=w.wind(k), damping=w.damping, spring_k=kk,
                                         sphere_radius=w.sphere_radius, friction=w.friction, hair_hair=True,
                                         grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h,
                                         colliders=np.stack(cols), seg_length=g.seg, wind_noise=g.noise,
                                         noise_dir=g.noise_dir, mesh=recs, mesh_margin=w.mesh_margin))
            ctx.step_batch(items)
            for j, g in enumerate(gs):
                grids[j] = oracle.step32_mt(ops[j], refs[j], g.N, grids[j], 8)
        for j, g in enumerate(gs):
            out, _ = gpus[j].read()
            o = out.reshape(-1, 3, g.N, 4).copy()
            r = refs[j].reshape(-1, 3, g.N, 4).copy()
            o[:, 2] = 0
            r[:, 2] = 0
            assert_bitwise(o, r, f"strands of groom {j}", g.N)
            assert np.array_equal(gpus[j].read_grid(), grids[j].reshape(-1))
    finally:
        ctx.set_batch_graphs(False)
        for gp in gpus:
            gp.destroy()
        ctx.mesh_destroy(mid)
