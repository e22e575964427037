"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must still reproduce them bit for bit on the CPU, and the HIP path
must reproduce them bit for bit on the GPU."""
import os

import numpy as np
import pytest

import oracle
from hairsim import scenes
from helpers import assert_bitwise, positions, velocities

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def plain_params(k, seg):
    wl = scenes.plain_1k()
    return wl, oracle.Params(dt=wl.dt, damping=wl.damping, sphere_radius=wl.sphere_radius,
                             colliders=scenes.sphere_collider(wl.sphere(k)), seg_length=seg)


def test_oracle_reproduces_plain_fixture():
    fx = load("plain_1k_subset.npz")
    st = fx["initial"].copy()
    seg = np.full(st.shape[0], np.float32(1.0) / np.float32(15), np.float32)
    for k in range(10):
        _, p = plain_params(k, seg)
        oracle.step32(p, st, 16)
        if f"step{k + 1}" in fx:
            assert_bitwise(st, fx[f"step{k + 1}"], f"step {k + 1}", 16)


def friction_params(k, fx):
    wl = scenes.plain_100k(S=64, N=16, G=16)
    g = wl.grooms[0]
    return wl, oracle.Params(dt=wl.dt, wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                             friction=wl.friction, hair_hair=True, grid_dim=16, grid_origin=g.grid_origin,
                             grid_h=g.grid_h, colliders=scenes.sphere_collider(wl.sphere(k)),
                             seg_length=fx["seg"], wind_noise=fx["noise"])


def test_oracle_reproduces_friction_fixture():
    fx = load("friction_64.npz")
    st = fx["initial"].copy()
    grid = None
    for k in range(3):
        _, p = friction_params(k, fx)
        grid = oracle.step32(p, st, 16, grid)
        if f"step{k + 1}" in fx:
            assert_bitwise(st, fx[f"step{k + 1}"], f"step {k + 1}", 16)
            assert np.array_equal(grid, fx[f"grid{k + 1}"])


@pytest.mark.gpu
def test_gpu_reproduces_fixtures(ctx):
    import hairsim

    fx = load("plain_1k_subset.npz")
    seg = np.full(32, np.float32(1.0) / np.float32(15), np.float32)
    gr = ctx.groom(fx["initial"], 16, seg_length=seg)
    wl = scenes.plain_1k()
    for k in range(10):
        gr.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))],
                hairsim.step_params(damping=wl.damping, sphere_radius=wl.sphere_radius, num_colliders=1,
                                    hair_hair=False))
        if f"step{k + 1}" in fx:
            out, _ = gr.read()
            ref = fx[f"step{k + 1}"]
            assert_bitwise(positions(out, 16), positions(ref, 16), "positions")
            assert_bitwise(velocities(out, 16), velocities(ref, 16), "velocities")
    gr.destroy()

    fx = load("friction_64.npz")
    wl = scenes.plain_100k(S=64, N=16, G=16)
    g = wl.grooms[0]
    gr = ctx.groom(fx["initial"], 16, seg_length=fx["seg"], wind_noise=fx["noise"], grid_dim=16,
                   grid_origin=g.grid_origin, grid_h=g.grid_h)
    for k in range(3):
        gr.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))],
                hairsim.step_params(wind=wl.wind(k), damping=wl.damping, sphere_radius=wl.sphere_radius,
                                    friction=wl.friction, num_colliders=1, hair_hair=True))
        if f"step{k + 1}" in fx:
            out, _ = gr.read()
            ref = fx[f"step{k + 1}"]
            assert_bitwise(positions(out, 16), positions(ref, 16), "positions")
            assert_bitwise(velocities(out, 16), velocities(ref, 16), "velocities")
            assert np.array_equal(gr.read_grid(), fx[f"grid{k + 1}"])
    gr.destroy()
