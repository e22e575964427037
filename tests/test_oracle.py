"""Pins of the CPU oracle (oracle/libhair_oracle.so) that do not restate its code.

The reference ships no tests, fixtures or golden vectors for the physics pass,
so each stage of compute.comp is pinned here by what the mathematics fixes:
closed forms, invariants, brute force on tiny inputs and relations between the
outputs, chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle
from hairsim import scenes

G = 64


def grid_params(**kw):
    args = dict(dt=1.0 / 60.0, hair_hair=True, grid_dim=G, grid_origin=(0.0, 0.0, 0.0), grid_h=1.0)
    args.update(kw)
    return oracle.Params(**args)


def cells_of(grid):
    g = grid.reshape(-1, 4)
    nz = np.flatnonzero(np.any(g != 0, axis=1))
    return {int(i): tuple(int(v) for v in g[i]) for i in nz}


# ---------------------------------------------------------------- int() --
@pytest.mark.parametrize("f,expect", [
    (2.9, 2), (-2.9, -2), (0.0, 0), (-0.0, 0), (float("nan"), 0),
    (float("inf"), 2**31 - 1), (-float("inf"), -2**31), (2147483648.0, 2**31 - 1),
    (-2147483648.0, -2**31), (2147483520.0, 2147483520), (-1e12, -2**31), (1e12, 2**31 - 1)])
def test_saturating_int(f, expect):
    assert oracle.sat_i32(f) == expect


# ---------------------------------------------------------------- splat --
def test_splat_single_point_dyadic_weights():
    """q = (10.25, 20.5, 30.75): trilinear weights are dyadic, so every int is exact."""
    p = grid_params()
    grid = p.grid_buffer()
    v = np.array([1.5, -2.25, 0.5])
    oracle.splat32(p, [10.25, 20.5, 30.75], v, grid)
    wx = {10: 0.75, 11: 0.25}
    wy = {20: 0.5, 21: 0.5}
    wz = {30: 0.25, 31: 0.75}
    expect = {}
    for a, wa in wx.items():
        for b, wb in wy.items():
            for c, wc in wz.items():
                w = wa * wb * wc
                expect[a + b * G + c * G * G] = (int(1e6 * w * v[0]), int(1e6 * w * v[1]), int(1e6 * w * v[2]),
                                                 int(1e6 * w))
    assert cells_of(grid) == expect
    assert sum(c[3] for c in expect.values()) == 1_000_000


def test_splat_edges():
    """Below 0 only cell 0 is touched (weight 1-|q|); at or beyond G nothing is; just below G only G-1."""
    p = grid_params()
    grid = p.grid_buffer()
    oracle.splat32(p, [-0.25, 5.0, 5.0], [0.0, 0.0, 0.0], grid)
    got = cells_of(grid)
    assert set(got) == {0 + 5 * G + 5 * G * G}
    assert got[5 * G + 5 * G * G][3] == 750000
    grid[:] = 0
    oracle.splat32(p, [G + 0.0, 5.0, 5.0], [1.0, 1.0, 1.0], grid)
    assert cells_of(grid) == {}
    grid[:] = 0
    oracle.splat32(p, [-1.0, 5.0, 5.0], [1.0, 1.0, 1.0], grid)
    assert cells_of(grid) == {}
    grid[:] = 0
    oracle.splat32(p, [G - 0.5, 5.0, 5.0], [0.0, 0.0, 0.0], grid)
    got = cells_of(grid)
    assert set(got) == {(G - 1) + 5 * G + 5 * G * G}
    assert got[(G - 1) + 5 * G + 5 * G * G][3] == 500000


def test_splat_conservation_random_points():
    """Interior points: density sums to 1e6 minus at most 8 truncations (+1 for fp rounding)."""
    rng = np.random.default_rng(3)
    p = grid_params()
    for _ in range(200):
        q = rng.uniform(1.0, G - 2.0, 3)
        v = rng.normal(0, 3, 3)
        grid = p.grid_buffer()
        oracle.splat32(p, q, v, grid)
        g = grid.reshape(-1, 4).astype(np.int64)
        d = g[:, 3].sum()
        assert 1_000_000 - 9 <= d <= 1_000_000
        mom = g[:, :3].sum(axis=0)
        assert np.all(np.abs(mom - 1e6 * v) <= 8 + 1e-5 * np.abs(1e6 * v))


def test_splat_order_independent_and_mt_matches_st():
    wl = scenes.plain_100k(S=3000, N=12, G=32)
    g = wl.grooms[0]

    def params():
        return oracle.Params(dt=wl.dt, wind=wl.wind(0), damping=wl.damping, sphere_radius=wl.sphere_radius,
                             friction=wl.friction, hair_hair=True, grid_dim=g.grid_dim, grid_origin=g.grid_origin,
                             grid_h=g.grid_h, colliders=scenes.sphere_collider(wl.sphere(0)), seg_length=g.seg,
                             wind_noise=g.noise)
    a = g.strands.copy()
    ga = oracle.step32(params(), a, g.N)
    b = g.strands.copy()
    gb = oracle.step32_mt(params(), b, g.N, threads=4)
    assert np.array_equal(ga, gb)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    perm = np.random.default_rng(1).permutation(g.S)
    c = np.ascontiguousarray(g.strands[perm])
    pc = oracle.Params(dt=wl.dt, wind=wl.wind(0), damping=wl.damping, sphere_radius=wl.sphere_radius,
                       friction=wl.friction, hair_hair=True, grid_dim=g.grid_dim, grid_origin=g.grid_origin,
                       grid_h=g.grid_h, colliders=scenes.sphere_collider(wl.sphere(0)), seg_length=g.seg[perm],
                       wind_noise=g.noise[perm])
    gc = oracle.step32(pc, c, g.N)
    assert np.array_equal(ga, gc)
    assert np.array_equal(a[perm].view(np.uint32), c.view(np.uint32))


# --------------------------------------------------------------- gather --
def test_gather_uniform_field_recovers_velocity():
    """Every voxel around the query occupied by points of one velocity v: the gather
    (sum over the 8 voxels of w * vel/density, weights summing to 1) returns v."""
    rng = np.random.default_rng(5)
    p = grid_params()
    grid = p.grid_buffer()
    v = np.array([0.7, -1.3, 2.1])
    for _ in range(3000):
        oracle.splat32(p, rng.uniform(20, 26, 3), v, grid)
    for _ in range(20):
        q = rng.uniform(21, 25, 3)
        gv = oracle.gather32(p, q, grid)
        assert np.allclose(gv, v, rtol=0, atol=2e-4), (gv, v)
        gv64 = oracle.gather64(p, q, grid)
        assert np.allclose(gv64, v, rtol=0, atol=2e-4)


def test_gather_single_point_and_skip_of_empty_cells():
    p = grid_params()
    grid = p.grid_buffer()
    v = np.array([1.0, 2.0, -3.0])
    oracle.splat32(p, [7.5, 7.5, 7.5], v, grid)
    assert np.allclose(oracle.gather32(p, [7.5, 7.5, 7.5], grid), v, atol=1e-5)
    # a cell with density <= 0 is skipped (compute.comp:281)
    grid[:] = 0
    grid[4 * (3 + 3 * G + 3 * G * G):][:4] = [1000, 1000, 1000, 0]
    assert np.array_equal(oracle.gather32(p, [3.0, 3.0, 3.0], grid), np.zeros(3, np.float32))
    grid[4 * (3 + 3 * G + 3 * G * G) + 3] = -5
    assert np.array_equal(oracle.gather32(p, [3.0, 3.0, 3.0], grid), np.zeros(3, np.float32))


# --------------------------------------------------------------- forces --
def sphere(center):
    return scenes.sphere_collider(center)


def test_force_gravity_wind_noise():
    p = oracle.Params(dt=1 / 60, gravity=(0.0, -9.8, 0.0), wind=(1.5, 0.0, -0.5), colliders=None,
                      wind_noise=np.zeros(1, np.float32), noise_dir=(0.0, 0.0, 2.0))
    hits, f = oracle.force64(p, [5.0, 5.0, 5.0], 0.25)
    assert hits == 0
    assert np.allclose(f, [1.5, -9.8, -0.5 + 0.5], atol=1e-6)


def test_force_sphere_closed_form():
    """Inside at depth r - delta: F = g + k (r - delta) n; outside: F = g (binary32-exact constants)."""
    c = np.array([0.0, 0.0, 1.25])
    g = np.array([0.0, -8.0, 0.0])
    p = oracle.Params(dt=1 / 64, gravity=tuple(g), sphere_radius=0.375, spring_k=1900.0, colliders=sphere(c))
    rng = np.random.default_rng(2)
    for _ in range(50):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        delta = rng.uniform(0.01, 0.37)
        hits, f = oracle.force64(p, c + delta * n, 0.0)
        assert hits == 1
        assert np.allclose(f, g + 1900.0 * (0.375 - delta) * n, rtol=1e-9, atol=1e-9)
        hits, f = oracle.force64(p, c + 0.38 * n, 0.0)
        assert hits == 0 and np.array_equal(f, g)


def ellipsoid(center, scale):
    T = np.diag([scale[0], scale[1], scale[2], 1.0])
    T[:3, 3] = center
    inv = np.linalg.inv(T)
    return np.concatenate([T.T.reshape(-1), inv.T.reshape(-1), inv.reshape(-1)]).astype(np.float32)


def test_force_ellipsoid_unit_scale_is_unit_sphere():
    """An ellipsoid collider with unit scale and no rotation acts as a unit sphere at its centre."""
    c = np.array([0.5, 2.0, -1.0])
    cols = np.stack([sphere([100.0, 100.0, 100.0]), ellipsoid(c, (1, 1, 1))])
    p = oracle.Params(dt=1 / 60, colliders=cols)
    rng = np.random.default_rng(4)
    for _ in range(30):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        delta = rng.uniform(0.05, 0.95)
        hits, f = oracle.force64(p, c + delta * n, 0.0)
        assert hits == 1
        assert np.allclose(f, np.array([0, -9.8, 0]) + 1900.0 * (1 - delta) * n, rtol=1e-6, atol=1e-6)


def test_force_ellipsoid_axis_scaled_and_average():
    c = np.array([0.0, 0.0, 0.0])
    cols = np.stack([sphere([0.0, 0.0, 0.0]), ellipsoid(c, (2, 1, 1))])
    p = oracle.Params(dt=1 / 64, gravity=(0.0, -8.0, 0.0), sphere_radius=1.5, colliders=cols)
    # on the x axis at 1: inside the ellipsoid (x/2 = 0.5); surface point (2,0,0): depth 1, normal +x.
    # the sphere (r 1.5) is hit too with depth 0.5, normal +x: force = g + (1900*1 + 1900*0.5)/2 x
    hits, f = oracle.force64(p, [1.0, 0.0, 0.0], 0.0)
    assert hits == 2
    assert np.allclose(f, [1900.0 * 1.5 / 2, -8.0, 0.0], rtol=1e-12)


# ------------------------------------------------------- mesh collider --
def _np_point_tri_dist(x, a, b, c):
    """Distance from x to every triangle (a, b, c: (T, 3)), float64, computed
    independently of the oracle: plane distance when the projection falls
    inside the triangle (barycentric test), else the nearest of the three
    edge segments."""
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    dp = np.einsum("ij,ij->i", x - a, n)
    proj = x - dp[:, None] * n
    v0, v1, v2 = b - a, c - a, proj - a
    d00, d01, d11 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v1 * v1).sum(1)
    d20, d21 = (v2 * v0).sum(1), (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)

    def seg(p0, p1):
        d = p1 - p0
        t = np.clip(np.einsum("ij,ij->i", x - p0, d) / (d * d).sum(1), 0.0, 1.0)
        return np.linalg.norm(x - (p0 + t[:, None] * d), axis=1)
    edge = np.minimum(np.minimum(seg(a, b), seg(b, c)), seg(c, a))
    return np.where(inside, np.abs(dp), edge)


def _mesh_params(V, T, margin):
    return oracle.Params(dt=1 / 64, gravity=(0.0, -8.0, 0.0), spring_k=1900.0, colliders=None,
                         mesh=oracle.mesh_prepare(V, T), mesh_margin=margin)


def test_mesh_face_centre_closed_form():
    """Just inside a face centre of a convex mesh: F = g + k d n_face, d the depth; just outside: F = g."""
    V, T = scenes.icosphere(2, (1.0, 1.2, 1.0))
    p = _mesh_params(V, T, 0.1)
    g = np.array([0.0, -8.0, 0.0])
    V64 = V.astype(np.float64)
    for f in range(0, len(T), 7):
        a, b, c = V64[T[f]]
        cen = (a + b + c) / 3
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        for depth in (0.003, 0.02):
            hits, F = oracle.force64(p, cen - depth * n, 0.0)
            assert hits == 1, (f, depth)
            assert np.allclose(F, g + 1900.0 * depth * n, rtol=1e-6, atol=1e-6 * 1900)
        hits, F = oracle.force64(p, cen + 0.01 * n, 0.0)
        assert hits == 0 and np.array_equal(F, g)
        hits, F = oracle.force64(p, cen - 0.15 * n, 0.0)  # deeper than the margin
        assert hits == 0


def test_mesh_contact_matches_brute_force_distance_and_convex_inside():
    """Random points near a scaled icosphere: contact iff inside (all face planes)
    and within the margin; depth = min point-triangle distance (independent numpy)."""
    V, T = scenes.icosphere(2, (1.0, 1.2, 1.0))
    margin = 0.08
    p = _mesh_params(V, T, margin)
    V64 = V.astype(np.float64)
    a, b, c = V64[T[:, 0]], V64[T[:, 1]], V64[T[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    rng = np.random.default_rng(11)
    g = np.array([0.0, -8.0, 0.0])
    seen = {0: 0, 1: 0}
    for _ in range(400):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = d * np.array([1.0, 1.2, 1.0]) * rng.uniform(0.88, 1.06)
        dist = _np_point_tri_dist(x, a, b, c).min()
        inside = (np.einsum("ij,ij->i", x - a, n) < 0).all()
        hits, F = oracle.force64(p, x, 0.0)
        if abs(dist - margin) < 1e-9 or dist < 1e-9:
            continue
        want = 1 if (inside and dist < margin) else 0
        assert hits == want, (x, dist, inside)
        seen[want] += 1
        if hits:
            assert abs(np.linalg.norm(F - g) - 1900.0 * dist) <= 1e-6 * 1900.0
    assert seen[0] > 50 and seen[1] > 50


def test_mesh_rejects_open_mesh():
    V, T = scenes.icosphere(1)
    with pytest.raises(ValueError):
        oracle.mesh_prepare(V, T[:-1])


# ----------------------------------------------------- predict + FTL ----
def test_single_segment_closed_form():
    """N = 2: the free point moves to the projection of its prediction on the sphere |p - root| = r."""
    dt, r, D = 0.125, 0.5, 0.875   # binary32-exact
    st = np.zeros((1, 24), np.float64)
    st[0, 4:7] = [r, 0.0, 0.0]
    p = oracle.Params(dt=dt, damping=D, gravity=(0.0, -8.0, 0.0), seg_length=np.array([r], np.float32))
    x = np.array([r, 0.0, 0.0])
    oracle.step64(p, st, 2)
    drop = dt * dt * 8.0
    th = math.atan2(drop, r)
    pn = np.array([r * math.cos(th), -r * math.sin(th), 0.0])
    assert np.allclose(st[0, 4:7], pn, rtol=0, atol=1e-15)
    assert np.allclose(st[0, 12:15], D * (pn - x) / dt, rtol=0, atol=1e-13)  # vel of point 1 (N=2)
    assert np.allclose(st[0, :3], 0.0) and st[0, 0:3].tolist() == [0.0, 0.0, 0.0]


def random_strands(rng, S, N, r):
    roots = rng.normal(size=(S, 3))
    st = np.zeros((S, 3, N, 4))
    for s in range(S):
        p = roots[s].copy()
        for i in range(N):
            st[s, 0, i, :3] = p
            st[s, 0, i, 3] = 1.0
            d = rng.normal(size=3)
            p = p + r * d / np.linalg.norm(d)
        st[s, 1, :, :3] = rng.normal(scale=0.5, size=(N, 3))
    return st.reshape(S, 12 * N)


def test_ftl_invariants_fp64_and_fp32():
    rng = np.random.default_rng(9)
    S, N, r = 64, 12, 0.07
    st = random_strands(rng, S, N, r)
    seg = np.full(S, r, np.float32)
    p = oracle.Params(dt=1 / 60, damping=0.9, seg_length=seg, sphere_radius=0.3, colliders=sphere([0, 0, 0]))
    before = st.copy()
    oracle.step64(p, st, N)
    P0, P1 = before.reshape(S, 3, N, 4)[:, 0, :, :3], st.reshape(S, 3, N, 4)[:, 0, :, :3]
    assert np.array_equal(P1[:, 0], P0[:, 0])
    seglen = np.linalg.norm(np.diff(P1, axis=1), axis=2)
    assert np.max(np.abs(seglen / np.float64(np.float32(r)) - 1)) < 1e-12
    # FP32: segment length within a few ulp of |p| relative to r
    st32 = before.astype(np.float32)
    oracle.step32(p, st32, N)
    P32 = st32.reshape(S, 3, N, 4)[:, 0, :, :3].astype(np.float64)
    seg32 = np.linalg.norm(np.diff(P32, axis=1), axis=2)
    ulp = np.spacing(np.float32(np.abs(P32).max())).astype(np.float64)
    assert np.max(np.abs(seg32 - np.float32(r))) <= 8 * ulp


def test_ftl_point_lies_on_ray_toward_prediction():
    rng = np.random.default_rng(11)
    S, N, r, dt = 32, 8, 0.125, 1 / 64
    st = random_strands(rng, S, N, r)
    seg = np.full(S, r, np.float32)
    p = oracle.Params(dt=dt, damping=0.875, gravity=(0.0, -8.0, 0.0), seg_length=seg)
    before = st.copy().reshape(S, 3, N, 4)
    oracle.step64(p, st, N)
    after = st.reshape(S, 3, N, 4)
    for s in range(S):
        for i in range(1, N):
            x, u = before[s, 0, i, :3], before[s, 1, i, :3]
            pred = x + dt * u + dt * dt * np.array([0, -8.0, 0])
            par = after[s, 0, i - 1, :3]
            d1 = after[s, 0, i, :3] - par
            d2 = pred - par
            assert np.linalg.norm(np.cross(d1, d2)) <= 1e-12 * np.linalg.norm(d2)
            assert np.dot(d1, d2) > 0


def test_dftl_relation_between_outputs():
    """Tip: v = D (p_new - x)/dt.  Others: v_i = D (p_new_i - x_i)/dt - c_{i+1}/dt, and
    c_i = D (p_new_i - pred_i) with pred_i = x_i + dt u_i + dt^2 F(x_i)."""
    rng = np.random.default_rng(13)
    S, N, r, dt, D = 16, 9, 0.05, 1 / 60, 0.9
    st = random_strands(rng, S, N, r) * 0.2
    seg = np.full(S, r, np.float32)
    p = oracle.Params(dt=dt, damping=D, vmax=1e9, seg_length=seg, wind=(0.5, 0.0, 0.2))
    before = st.copy().reshape(S, 3, N, 4)
    oracle.step64(p, st, N)
    after = st.reshape(S, 3, N, 4)
    inv_dt = 1.0 / np.float64(np.float32(dt))
    dtf = np.float64(np.float32(dt))
    for s in range(S):
        X, U = before[s, 0, :, :3], before[s, 1, :, :3]
        P, V, Cc = after[s, 0, :, :3], after[s, 1, :, :3], after[s, 2, :, :3]
        for i in range(1, N):
            pred = X[i] + dtf * U[i] + dtf * dtf * np.array([0.5, np.float32(-9.8), 0.2])
            assert np.allclose(Cc[i], D * (P[i] - pred), atol=1e-12)
            base = np.float64(np.float32(D)) * (P[i] - X[i]) * inv_dt
            if i == N - 1:
                assert np.allclose(V[i], base, atol=1e-9)
            else:
                assert np.allclose(V[i], base - Cc[i + 1] * inv_dt, atol=1e-9)
        assert np.array_equal(V[0], U[0])


def test_velocity_clamp_boundary():
    """|v| just above vmax is clamped to vmax along v; just below is untouched."""
    dt, r, D = 1.0, 16.0, 1.0
    for target, clamped in ((10.000001, True), (9.999999, False)):
        # root at 0, point at (r,0,0); the prediction points at angle ang so that the FTL
        # displacement is a chord of length target (|v| = target with dt = D = 1)
        st = np.zeros((1, 24))
        st[0, 4:7] = [r, 0.0, 0.0]
        ang = 2 * math.asin(target / (2 * r))
        pred_dir = np.array([math.cos(ang), 0.0, math.sin(ang)])
        u = pred_dir * 20.0 - np.array([r, 0.0, 0.0])
        st[0, 12:15] = u  # vel of point 1 (N=2: floats 4N + 4)
        p = oracle.Params(dt=dt, damping=D, gravity=(0.0, 0.0, 0.0), seg_length=np.array([r], np.float32))
        oracle.step64(p, st, 2)
        v = st[0, 12:15]
        if clamped:
            assert abs(np.linalg.norm(v) - 10.0) < 1e-12
        else:
            assert abs(np.linalg.norm(v) - target) < 1e-9


def test_hanging_strand_positions_are_fixed():
    """A straight strand hanging along -g with no colliders keeps its positions (FP64)."""
    N, r = 10, float(np.float32(0.1))
    st = np.zeros((1, 3, N, 4))
    for i in range(N):
        st[0, 0, i] = [0.3, 2.0 - i * r, -0.2, 1.0]
    st = st.reshape(1, 12 * N)
    p = oracle.Params(dt=1 / 60, damping=0.9, seg_length=np.array([r], np.float32))
    before = st.copy()
    for _ in range(5):
        oracle.step64(p, st, N)
    P0 = before.reshape(1, 3, N, 4)[0, 0, :, :3]
    P1 = st.reshape(1, 3, N, 4)[0, 0, :, :3]
    assert np.max(np.abs(P1 - P0)) < 1e-12


# ---------------------------------------------------- fp32 vs fp64 ------
def test_ref32_tracks_oracle64_per_step():
    """Same FP32 state into both each step: positions within 1e-5 * max(|p|, 1)."""
    wl = scenes.plain_1k(S=256)
    g = wl.grooms[0]
    s32 = g.strands.copy()
    worst = 0.0
    for k in range(240):
        p = oracle.Params(dt=wl.dt, damping=wl.damping, sphere_radius=wl.sphere_radius,
                          colliders=sphere(wl.sphere(k)), seg_length=g.seg)
        s64 = s32.astype(np.float64)
        oracle.step64(p, s64, g.N)
        oracle.step32(p, s32, g.N)
        P32 = s32.reshape(-1, 3, g.N, 4)[:, 0, :, :3]
        P64 = s64.reshape(-1, 3, g.N, 4)[:, 0, :, :3]
        worst = max(worst, float(np.max(np.abs(P32 - P64) / np.maximum(np.abs(P64), 1.0))))
    assert worst <= 1e-5, worst


# ------------------------------------------- forcing (compute.comp:150-151) --
def _trig_samples(lo, hi, n, seed):
    return np.random.default_rng(seed).uniform(lo, hi, n).astype(np.float32)


@pytest.mark.parametrize("lo,hi,abs_tol,ulp_tol", [
    (-math.pi / 4, math.pi / 4, 8e-8, 1.5), (-10.0, 10.0, 1e-7, 2.0), (-1e5, 1e5, 1.2e-7, None)])
def test_forcing_sin_cos_against_binary64(lo, hi, abs_tol, ulp_tol):
    """The oracle's binary32 sin/cos against libm's binary64 ones: absolute
    error below 1.2e-7 everywhere up to 1e5, within 2 ulp near the origin."""
    L = oracle.lib()
    x = _trig_samples(lo, hi, 20000, 7)
    for f, ref in ((L.ho_sinf, np.sin), (L.ho_cosf, np.cos)):
        got = np.array([f(float(v)) for v in x], dtype=np.float64)
        want = ref(x.astype(np.float64))
        err = np.abs(got - want)
        assert err.max() < abs_tol, (f.__name__, err.max())
        if ulp_tol is not None:
            ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
            assert (err / ulp).max() <= ulp_tol, (f.__name__, (err / ulp).max())


def test_forcing_sin_cos_special_values():
    L = oracle.lib()
    assert L.ho_sinf(0.0) == 0.0 and L.ho_cosf(0.0) == 1.0
    assert math.isnan(L.ho_sinf(float("nan"))) and math.isnan(L.ho_cosf(float("inf")))
    # quadrant logic: sin(x + pi/2) = cos(x), sin(x + pi) = -sin(x), within
    # the rounding of the shifted argument to binary32
    for x in (0.3, -1.1, 2.0, 40.0):
        tol = 3e-7 + float(np.spacing(np.float32(abs(x) + 4.0)))
        assert abs(L.ho_sinf(x + math.pi / 2) - L.ho_cosf(x)) < tol
        assert abs(L.ho_sinf(x + math.pi) + L.ho_sinf(x)) < tol


def test_fbm_on_the_lattice_reduces_to_the_hash():
    """At integer points every octave sits on a lattice node, where noise()
    is the corner hash itself (u = 0): fbm = sum 2^-(i+1) random(2^i p)."""
    L = oracle.lib()
    f32 = np.float32
    for px, py in ((0.0, 0.0), (3.0, -2.0), (-7.0, 5.0), (1.0, 1.0)):
        want = f32(0.0)
        amp = f32(0.5)
        qx, qy = f32(px), f32(py)
        for _ in range(6):
            d = qx * f32(12.9898) + qy * f32(78.233)
            h = f32(L.ho_sinf(float(d))) * f32(43758.5453123)
            want = want + amp * (h - np.floor(h))
            qx, qy, amp = qx * f32(2), qy * f32(2), amp * f32(0.5)
        assert L.ho_fbmf(px, py) == want


def test_fbm_is_continuous_and_bounded():
    """The Hermite blend makes noise continuous across cell edges; a wrong
    corner or blend term would jump there.  Six octaves of [0,1) noise with
    amplitudes 1/2 .. 1/64 stay in [0, 63/64)."""
    L = oracle.lib()
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = float(rng.integers(-20, 20))
        y = float(rng.uniform(-20, 20))
        for a, b in (((n - 1e-4, y), (n + 1e-4, y)), ((y, n - 1e-4), (y, n + 1e-4))):
            assert abs(L.ho_fbmf(*a) - L.ho_fbmf(*b)) < 0.01
    vals = [L.ho_fbmf(float(x), float(y)) for x, y in rng.uniform(-50, 50, (2000, 2))]
    assert min(vals) >= 0.0 and max(vals) < 63.0 / 64.0


@pytest.mark.parametrize("mode", [1, 2])
def test_forcing_force_closed_form(mode):
    """Force with the forcing on, against the shader lines evaluated in
    binary64 (with the oracle's fbm value for mode 2): gravity plus the line,
    no colliders."""
    L = oracle.lib()
    rng = np.random.default_rng(20 + mode)
    for _ in range(50):
        t = float(rng.uniform(0, 30))
        x = rng.uniform(-3, 3, 3).astype(np.float32)
        p = oracle.Params(dt=1 / 60, total_time=t, forcing=mode)
        hits, f = oracle.force32(p, x)
        assert hits == 0
        X, Y, Z = (float(v) for v in x)
        tf = float(np.float32(t))
        sx = 2.0 * math.sin(tf * 2.0) * math.cos(Y * 10.0) * math.sin((Y + 5.0) * 15.0)
        if mode == 1:
            want = np.array([10 * sx, -9.8, -10 * min(max(Y * 2.0, 0.2), 2.0)])
        else:
            amp = 7.0 * L.ho_fbmf(L.ho_sinf(tf), L.ho_cosf(tf))
            want = np.array([amp * sx, -9.8 + amp * 4.0 * math.sin(Z * 5.0 + tf * 3.0), amp * -0.6 * (Y + 3.0)])
        np.testing.assert_allclose(f, want, rtol=2e-5, atol=2e-4)


def test_forcing_off_is_the_reference_force():
    x = np.array([0.4, 1.3, -0.2], dtype=np.float32)
    a = oracle.force32(oracle.Params(dt=1 / 60, total_time=3.0, forcing=0), x)[1]
    assert np.array_equal(a, np.array([0.0, -9.8, 0.0], dtype=np.float32))


# ------------------------------- curve resampling (hair.tese:34-80, func) --
def _strands_from_points(P):
    """(S, N, 3) positions -> Strand AoS (velocities, corrections zero)."""
    S, N, _ = P.shape
    a = np.zeros((S, 3, N, 4), dtype=np.float32)
    a[:, 0, :, :3] = P
    a[:, 0, :, 3] = 1.0
    return a.reshape(-1)


def _catmull_rom64(P, M):
    """Independent binary64 Hermite form of the same curve (tangents
    (v2 - v0)/2 and (v3 - v1)/2, mirrored end neighbours), same sample
    parameters v = j/(M-1) rounded to binary32."""
    S, N, _ = P.shape
    P = P.astype(np.float64)
    out = np.zeros((S, M, 3))
    for j in range(M):
        v = float(np.float32(j) / np.float32(M - 1))
        x = float(np.float32(v) * np.float32(N - 1))
        seg = int(math.floor(x))
        t = x - seg
        if seg >= N - 1:
            seg, t = N - 2, 1.0
        v1, v2 = P[:, seg], P[:, seg + 1]
        v0 = 2 * v1 - v2 if seg == 0 else P[:, seg - 1]
        v3 = 2 * v2 - v1 if seg + 2 == N else P[:, seg + 2]
        m1, m2 = (v2 - v0) / 2, (v3 - v1) / 2
        h = [2 * t**3 - 3 * t**2 + 1, t**3 - 2 * t**2 + t, -2 * t**3 + 3 * t**2, t**3 - t**2]
        out[:, j] = h[0] * v1 + h[1] * m1 + h[2] * v2 + h[3] * m2
    return out


def test_resample_straight_strand_is_uniform_line():
    """Equally spaced collinear points: the control points are equally spaced
    too, so the curve is the line at uniform speed (a wrong 1/3 or 1/2 factor
    or a wrong end mirror bends the speed)."""
    N, M = 9, 41
    d = np.array([0.3, -0.2, 0.1])
    P = (np.array([1.0, 2.0, -1.0]) + np.arange(N)[:, None] * d)[None].astype(np.float32)
    out = oracle.resample32(_strands_from_points(P), N, M)
    v = np.arange(M) / (M - 1)
    want = P[0, 0] + v[:, None] * (N - 1) * d
    np.testing.assert_allclose(out[0, :, :3], want, atol=2e-6)
    assert np.all(out[0, :, 3] == 1.0)


def test_resample_ends_are_the_strand_ends_and_nodes_interpolate():
    rng = np.random.default_rng(5)
    N = 12
    M = 3 * (N - 1) + 1
    P = np.cumsum(rng.normal(0, 0.1, (4, N, 3)), axis=1).astype(np.float32)
    out = oracle.resample32(_strands_from_points(P), N, M)
    assert np.array_equal(out[:, 0, :3], P[:, 0]) and np.array_equal(out[:, -1, :3], P[:, -1])
    np.testing.assert_allclose(out[:, ::3, :3], P, atol=1e-6)


@pytest.mark.parametrize("N,M", [(2, 5), (3, 2), (16, 64), (32, 7)])
def test_resample_matches_hermite_catmull_rom(N, M):
    rng = np.random.default_rng(N * 100 + M)
    P = np.cumsum(rng.normal(0, 0.2, (6, N, 3)), axis=1).astype(np.float32)
    out = oracle.resample32(_strands_from_points(P), N, M)
    np.testing.assert_allclose(out[:, :, :3], _catmull_rom64(P, M), atol=3e-6)
