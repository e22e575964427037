"""Shared helpers: drive the same workload through the HIP library and the oracle."""
from __future__ import annotations

import numpy as np

import hairsim
import oracle
from hairsim import scenes


def gpu_params(wl, k, groom=0, ncol=1, **kw):
    args = dict(wind=wl.wind(k, groom), damping=wl.damping, sphere_radius=wl.sphere_radius,
                friction=wl.friction, num_colliders=ncol, hair_hair=wl.hair_hair)
    args.update(kw)
    return hairsim.step_params(**args)


def oracle_params(wl, k, groom=0, colliders=None, **kw):
    g = wl.grooms[groom]
    if colliders is None:
        colliders = scenes.sphere_collider(wl.sphere(k, groom))
    args = dict(dt=wl.dt, wind=wl.wind(k, groom), damping=wl.damping, sphere_radius=wl.sphere_radius,
                friction=wl.friction, hair_hair=wl.hair_hair, grid_dim=g.grid_dim, grid_origin=g.grid_origin,
                grid_h=g.grid_h, colliders=colliders, seg_length=g.seg, wind_noise=g.noise,
                noise_dir=g.noise_dir)
    args.update(kw)
    return oracle.Params(**args)


def make_gpu_groom(ctx, wl, groom=0, keep_corrections=False):
    g = wl.grooms[groom]
    return ctx.groom(g.strands, g.N, seg_length=g.seg, wind_noise=g.noise, noise_dir=g.noise_dir,
                     grid_dim=g.grid_dim, grid_origin=g.grid_origin, grid_h=g.grid_h,
                     keep_corrections=keep_corrections)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1).view(np.uint32)


def assert_bitwise(gpu, ref, what="strands", N=None):
    gb, rb = bits(gpu), bits(ref)
    if not np.array_equal(gb, rb):
        bad = np.flatnonzero(gb != rb)
        msg = f"{what}: {bad.size} values differ; first index {bad[0]}"
        if N:
            s, r = divmod(int(bad[0]), 12 * N)
            arr, rem = divmod(r, 4 * N)
            msg += f" (strand {s}, {['pos', 'vel', 'corr'][arr]} point {rem // 4} comp {rem % 4})"
        msg += f": gpu {gpu.reshape(-1)[bad[0]]!r} ref {ref.reshape(-1)[bad[0]]!r}"
        raise AssertionError(msg)


def positions(aos, N):
    return aos.reshape(-1, 3, N, 4)[:, 0, :, :3]


def velocities(aos, N):
    return aos.reshape(-1, 3, N, 4)[:, 1, :, :3]


def run_parity(ctx, wl, steps, check_every=1, groom=0, keep_corrections=False, ncol=1, colliders_fn=None,
               params_kw=None, check_grid=True, tile_fn=None):
    """Step GPU and ref32 side by side; assert bitwise equality of strands (and grid).
    tile_fn(k): the tile (set_tile arguments) to step k with, or None to keep the current one."""
    g = wl.grooms[groom]
    gpu = make_gpu_groom(ctx, wl, groom, keep_corrections=keep_corrections)
    ref = g.strands.copy()
    grid = None
    params_kw = params_kw or {}
    try:
        for k in range(steps):
            if tile_fn is not None and tile_fn(k) is not None:
                ctx.set_tile(*tile_fn(k))
            cols = colliders_fn(k) if colliders_fn else [scenes.sphere_collider(wl.sphere(k, groom))]
            gpu.step(wl.dt, float(wl.time(k)), cols, gpu_params(wl, k, groom, ncol=ncol, **params_kw))
            op = oracle_params(wl, k, groom, colliders=np.stack(cols)[:ncol] if ncol else None, **params_kw)
            grid = oracle.step32(op, ref, g.N, grid)
            if (k + 1) % check_every == 0 or k == steps - 1:
                out, ind = gpu.read()
                if not keep_corrections:
                    # corrections are not kept on the device: compare pos + vel only
                    o = out.reshape(-1, 3, g.N, 4).copy()
                    r = ref.reshape(-1, 3, g.N, 4).copy()
                    o[:, 2] = 0
                    r[:, 2] = 0
                    assert_bitwise(o, r, f"step {k}", g.N)
                else:
                    assert_bitwise(out, ref, f"step {k}", g.N)
                assert ind == (g.S, 1, 0, 0)
                if check_grid and wl.hair_hair and g.grid_dim:
                    gg = gpu.read_grid()
                    if not np.array_equal(gg, grid):
                        bad = np.flatnonzero(gg != grid)
                        raise AssertionError(f"grid step {k}: {bad.size} ints differ, first cell {bad[0] // 4}")
    finally:
        gpu.destroy()
    return ref, grid
