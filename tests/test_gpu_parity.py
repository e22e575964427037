"""GPU parity: the HIP step against the oracle's FP32 restatement (bit for bit)
and its FP64 restatement (tolerance), through the C ABI.

Tolerances (stated here, checked below):
  * GPU vs ref32: bitwise equal strands and bitwise equal voxel ints.
  * GPU vs oracle64, one step from the same FP32 state: positions within
    1e-5 * max(|p|, 1); velocities within 1e-5 * max(|p|, 1) * damping / dt
    (the position tolerance carried through v = D (p_new - x) / dt).
"""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from helpers import (assert_bitwise, gpu_params, make_gpu_groom, oracle_params, positions, run_parity,
                     velocities)

pytestmark = pytest.mark.gpu


def test_plain_1k_240_steps_bitwise(ctx):
    """1024 x 16, sphere, no hair-hair: every step bitwise equal to ref32 for 240 steps."""
    wl = scenes.plain_1k()
    run_parity(ctx, wl, 240, check_every=1)


def test_plain_1k_keep_corrections_bitwise(ctx):
    wl = scenes.plain_1k(S=256)
    run_parity(ctx, wl, 12, keep_corrections=True)


def test_hair_hair_grid_bitwise(ctx):
    """Splat + global barrier + gather: strands and every voxel int bitwise."""
    wl = scenes.plain_100k(S=2048, N=16, G=32)
    run_parity(ctx, wl, 30, check_every=3)


def test_plain_100k_shape_reduced_bitwise(ctx):
    """The 100k x 32 workload's physics (wind, noise, orbiting sphere, 64^3 grid) on 8192 strands."""
    wl = scenes.plain_100k(S=8192, N=32, G=64)
    run_parity(ctx, wl, 8, check_every=2)


@pytest.mark.parametrize("S,N", [(1, 2), (63, 3), (65, 10), (1000, 64), (129, 7)])
def test_ragged_shapes_bitwise(ctx, S, N):
    wl = scenes.plain_100k(S=S, N=N, G=16)
    run_parity(ctx, wl, 5, check_every=1)


def reference_demo_colliders():
    """The six colliders of the reference demo (src/main.cpp:229-237)."""
    spec = [((2.0, 0.0, 1.0), (0, 0, 0), (1, 1, 1)),
            ((0.0, 2.64, 0.08), (-38.270, 0, 0), (0.817, 1.158, 1.01)),
            ((0.0, 1.35, -0.288), (18.301, 0, 0), (0.457, 1.0, 0.538)),
            ((0.0, -0.380, -0.116), (-17.260, 0, 0), (1.078, 1.683, 0.974)),
            ((-0.698, 0.087, -0.36), (-20.254, 13.144, 34.5), (0.721, 1.0, 0.724)),
            ((0.698, 0.087, -0.36), (-20.254, 13.144, -34.5), (0.721, 1.0, 0.724))]
    return [hairsim.collider_from_trs(*s).as_array() for s in spec]


def test_reference_constants_six_colliders_bitwise(ctx):
    """Reference constants (DAMPING 0.998, friction 0.08, 64^3 grid at (-3,-2,-5), h 7/64, sphere r 1,
    five ellipsoids) on strands hanging through the demo colliders."""
    rng = np.random.default_rng(7)
    S, N = 900, 10
    roots, nrm = scenes.hemisphere_roots(rng, S, radius=1.0)
    roots = roots + np.array([0.0, 2.3, 0.0])
    aos, seg = scenes.straight_strands(roots, nrm, N, np.full(S, 2.5))
    g = scenes.Groom(aos, N, seg, None, grid_dim=64, grid_origin=(-3.0, -2.0, -5.0), grid_h=7.0 / 64.0)
    wl = scenes.Workload("reference-demo", dt=1.0 / 60.0, steps=20, damping=0.998, friction=0.08,
                         hair_hair=True, sphere_radius=1.0, grooms=[g])
    cols = reference_demo_colliders()
    run_parity(ctx, wl, 20, check_every=2, ncol=6, colliders_fn=lambda k: cols)


def test_vs_oracle64_per_step_tolerance(ctx):
    """One step at a time from the GPU's FP32 state: GPU vs oracle64 within the stated tolerance."""
    wl = scenes.plain_100k(S=2048, N=32, G=32)
    g = wl.grooms[0]
    gpu = make_gpu_groom(ctx, wl)
    worst_p = worst_v = 0.0
    try:
        for k in range(30):
            before, _ = gpu.read()
            c = scenes.sphere_collider(wl.sphere(k))
            gpu.step(wl.dt, float(wl.time(k)), [c], gpu_params(wl, k))
            after, _ = gpu.read()
            s64 = before.astype(np.float64)
            oracle.step64(oracle_params(wl, k), s64, g.N)
            p64, p32 = positions(s64, g.N), positions(after, g.N)
            v64, v32 = velocities(s64, g.N), velocities(after, g.N)
            scale = np.maximum(np.abs(p64), 1.0)
            ep = np.max(np.abs(p32 - p64) / scale)
            ev = np.max(np.abs(v32 - v64) / (scale * wl.damping / wl.dt))
            worst_p, worst_v = max(worst_p, ep), max(worst_v, ev)
    finally:
        gpu.destroy()
    assert worst_p <= 1e-5, worst_p
    assert worst_v <= 1e-5, worst_v


def test_full_100k_properties(ctx):
    """Full 100,000 x 32 workload, 20 steps: finite, roots fixed, segment lengths, determinism."""
    wl = scenes.plain_100k()
    g = wl.grooms[0]
    outs = []
    for rep in range(2):
        gpu = make_gpu_groom(ctx, wl)
        for k in range(20):
            gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
        out, ind = gpu.read()
        grid = gpu.read_grid()
        gpu.destroy()
        outs.append((out, grid))
        assert ind == (g.S, 1, 0, 0)
    (a, ga), (b, gb) = outs
    assert_bitwise(a, b, "run-to-run determinism", g.N)
    assert np.array_equal(ga, gb)
    P = positions(a, g.N)
    assert np.isfinite(a).all()
    assert np.array_equal(P[:, 0], positions(g.strands, g.N)[:, 0])
    seg = np.linalg.norm(np.diff(P.astype(np.float64), axis=1), axis=2)
    rel = np.abs(seg / g.seg[:, None].astype(np.float64) - 1.0)
    # FP32 bound: endpoint rounding of |p| <= 4 against r >= 0.8/31
    ulp = np.spacing(np.float32(4.0)).astype(np.float64)
    assert rel.max() <= 8 * ulp / (0.8 / 31), rel.max()
    # every point lands in the grid: total density ~ 1e6 per point
    dens = ga.reshape(-1, 4)[:, 3].astype(np.int64).sum()
    npts = g.S * (g.N - 1)
    assert npts * (1e6 - 8) <= dens <= npts * 1e6


def test_full_1m_head_mesh_properties(ctx):
    """Full 1,000,000 x 64 workload with the head mesh and a 128^3 grid, 6
    steps, twice: the two runs are bit-identical (integer voxel sums and the
    mesh pass are order-independent), every value is finite, roots stay, the
    segments keep their rest length, and the density sums to about 1e6 per
    point (0.2 % of the points start beyond the 128^3 grid's edge)."""
    import hashlib
    wl = scenes.large_1m()
    g = wl.grooms[0]
    digests = []
    out = grid = None
    for rep in range(2):
        mid = ctx.mesh(*wl.mesh, grid_dim=wl.mesh_grid, margin=wl.mesh_margin)
        gpu = make_gpu_groom(ctx, wl)
        for k in range(6):
            gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))],
                     gpu_params(wl, k, mesh_id=mid))
        out, ind = gpu.read()
        grid = gpu.read_grid()
        gpu.destroy()
        ctx.mesh_destroy(mid)
        assert ind == (g.S, 1, 0, 0)
        digests.append(hashlib.sha1(out.tobytes() + grid.tobytes()).hexdigest())
    assert digests[0] == digests[1]
    assert np.isfinite(out).all()
    P = positions(out, g.N)
    assert np.array_equal(P[:, 0], positions(g.strands, g.N)[:, 0])
    seg = np.linalg.norm(np.diff(P, axis=1).astype(np.float64), axis=2)
    rel = np.abs(seg / g.seg[:, None].astype(np.float64) - 1.0)
    ulp = np.spacing(np.float32(4.0)).astype(np.float64)
    assert rel.max() <= 8 * ulp / float(g.seg.min()), rel.max()
    dens = grid.reshape(-1, 4)[:, 3].astype(np.int64).sum()
    npts = g.S * (g.N - 1)
    assert 0.99 * npts * (1e6 - 8) <= dens <= npts * 1e6


def test_read_is_non_destructive_and_write_roundtrip(ctx):
    wl = scenes.plain_100k(S=300, N=8, G=16)
    g = wl.grooms[0]
    gpu = make_gpu_groom(ctx, wl, keep_corrections=True)
    try:
        for k in range(3):
            gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
        a, _ = gpu.read()
        b, _ = gpu.read()
        assert_bitwise(a, b, "second read")
        gpu.write(a)
        c, _ = gpu.read()
        assert_bitwise(a, c, "write/read round trip")
    finally:
        gpu.destroy()


def test_hair_hair_toggle_and_friction_zero(ctx):
    """Grid off / on / off across steps, and friction 0 with the grid on equals the grid off."""
    wl = scenes.plain_100k(S=512, N=16, G=16)
    run_parity(ctx, wl, 6, params_kw=dict(hair_hair=False))
    wl2 = scenes.plain_100k(S=512, N=16, G=16)
    wl2.friction = 0.0
    r_on, _ = run_parity(ctx, wl2, 6)
    wl3 = scenes.plain_100k(S=512, N=16, G=16)
    r_off, _ = run_parity(ctx, wl3, 6, params_kw=dict(hair_hair=False))
    assert np.array_equal(positions(r_on, 16), positions(r_off, 16))
    assert np.array_equal(velocities(r_on, 16), velocities(r_off, 16))


def test_errors_are_status_codes(ctx):
    wl = scenes.plain_1k(S=64)
    gpu = make_gpu_groom(ctx, wl)
    with pytest.raises(hairsim.HairSimError) as e:
        gpu.step(0.0, 0.0, [], gpu_params(wl, 0, ncol=0))
    assert e.value.code == hairsim.HS_ERR_INVALID
    with pytest.raises(hairsim.HairSimError) as e:
        gpu.step(wl.dt, 0.0, [], gpu_params(wl, 0, ncol=2))
    assert e.value.code == hairsim.HS_ERR_INVALID
    gpu.destroy()
    gpu.id = 12345
    with pytest.raises(hairsim.HairSimError) as e:
        gpu.read()
    assert e.value.code == hairsim.HS_ERR_NOT_FOUND


def _stacked_strands(S, root, drop):
    """S identical two-point strands: root, and a point `drop` straight below it, at rest."""
    aos = np.zeros((S, 3, 2, 4), dtype=np.float32)
    aos[:, 0, 0, :3] = root
    aos[:, 0, 1, :3] = (root[0], root[1] - drop, root[2])
    aos[:, 0, :, 3] = 1.0
    return aos.reshape(-1)


@pytest.mark.parametrize("S", [1100, 76800])
def test_stacked_voxel_near_limit_and_list_overflow(ctx, S):
    """Every strand's moving point sits on one voxel corner (weight 1 in one cell).

    S = 1100: density 1.1e9 >= 2^30 in one cell -> one near-limit cell per step.
    S = 76800: 600 workgroup boxes overlap one brick, more than a brick's box
    list holds, so the reduce falls back to scanning every box; the density
    wraps in int32 exactly as the shader's atomicAdd does.
    """
    G, h, origin = 16, 0.25, (-2.0, -2.0, -2.0)
    root, drop = (0.0, 0.5, 0.0), 0.5
    aos = _stacked_strands(S, root, drop)
    seg = np.full(S, drop, dtype=np.float32)
    gpu = ctx.groom(aos, 2, seg_length=seg, grid_dim=G, grid_origin=origin, grid_h=h)
    ref = aos.copy()
    grid = None
    try:
        for k in range(3):
            p = hairsim.step_params(damping=0.9, num_colliders=0, friction=0.05)
            gpu.step(1.0 / 64, (k + 1) / 64, [], p)
            op = oracle.Params(dt=1.0 / 64, damping=0.9, friction=0.05, hair_hair=True, grid_dim=G,
                               grid_origin=origin, grid_h=h, seg_length=seg)
            grid = oracle.step32(op, ref, 2, grid)
            out, _ = gpu.read()
            o = out.reshape(-1, 3, 2, 4).copy()
            r = ref.reshape(-1, 3, 2, 4).copy()
            o[:, 2] = 0
            r[:, 2] = 0
            assert_bitwise(o, r, f"strands after step {k}", 2)
            gg = gpu.read_grid()
            assert np.array_equal(gg, grid.reshape(-1)), f"grid after step {k}"
            near = int((np.abs(gg.astype(np.int64)) >= 2 ** 30).reshape(-1, 4).any(axis=1).sum())
            if S == 1100:
                assert near == 1
            assert gpu.grid_near_limit() == (k + 1) * near
    finally:
        gpu.destroy()


def test_reference_demo_groom_bitwise(ctx):
    """The reference demo end to end: its 900 roots on the mannequin mesh
    (tests/golden/mannequin_roots.npz, srand(8) over the triangulated OBJ),
    Hair::Hair's 10-point strands (hs_initial_strands), the six demo colliders
    and the reference constants; 30 steps bitwise vs ref32."""
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "mannequin_roots.npz"))
    N = 10
    S = fx["roots"].shape[0]
    aos = hairsim.initial_strands(fx["roots"], fx["normals"], N).reshape(S, 12 * N)
    g = scenes.Groom(aos, N, np.full(S, np.float32(2.5 / 9.0)), None, grid_dim=64, grid_origin=(-3.0, -2.0, -5.0),
                     grid_h=7.0 / 64.0)
    wl = scenes.Workload("reference-demo-mannequin", dt=1.0 / 60.0, steps=30, damping=0.998, friction=0.08,
                         hair_hair=True, sphere_radius=1.0, grooms=[g])
    cols = reference_demo_colliders()
    run_parity(ctx, wl, 30, check_every=5, ncol=6, colliders_fn=lambda k: cols)
