"""Render-facing output on the device: hs_export_strands (the strand SSBO +
indirect record the reference renderer draws from) against the host read-back,
and hs_resample_strands (the tessellation shader's curve, hair.tese:34-80)
against the oracle's restatement, bit for bit."""
import numpy as np
import pytest

import hairsim
import oracle
from hairsim import scenes
from helpers import assert_bitwise, gpu_params, make_gpu_groom

pytestmark = pytest.mark.gpu


def _stepped(ctx, wl, steps):
    gpu = make_gpu_groom(ctx, wl)
    for k in range(steps):
        gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
    return gpu


def test_export_equals_read_back(ctx):
    """Device export of a groom with pending grid friction: the same bytes as
    hs_read_strands, and the indirect record {S, 1, 0, 0}."""
    wl = scenes.plain_100k(S=1000, N=16, G=32)
    gpu = _stepped(ctx, wl, 6)
    S, N = wl.grooms[0].S, wl.grooms[0].N
    buf = hairsim.DeviceBuffer(48 * S * N)
    ind = hairsim.DeviceBuffer(16)
    try:
        buf.fill(0xFF)
        ind.fill(0xFF)
        gpu.export(buf, ind)
        ctx.sync()
        host, rec = gpu.read()
        assert_bitwise(buf.numpy(np.float32), host, "export", N)
        assert tuple(ind.numpy(np.uint32)) == (S, 1, 0, 0) == rec
    finally:
        buf.free()
        ind.free()
        gpu.destroy()


def test_export_with_kept_corrections(ctx):
    """A groom that keeps its correction vectors exports them too."""
    wl = scenes.plain_100k(S=300, N=12, G=16)
    g = wl.grooms[0]
    gpu = make_gpu_groom(ctx, wl, keep_corrections=True)
    for k in range(4):
        gpu.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
    buf = hairsim.DeviceBuffer(48 * g.S * g.N)
    try:
        gpu.export(buf)
        ctx.sync()
        host, _ = gpu.read()
        out = buf.numpy(np.float32)
        assert_bitwise(out, host, "export with corrections", g.N)
        assert np.any(out.reshape(-1, 3, g.N, 4)[:, 2] != 0)
    finally:
        buf.free()
        gpu.destroy()


@pytest.mark.parametrize("S,N,M", [(1000, 16, 64), (130, 2, 5), (257, 32, 7), (64, 10, 2)])
def test_resample_bitwise(ctx, S, N, M):
    wl = scenes.plain_100k(S=S, N=N, G=16)
    gpu = _stepped(ctx, wl, 5)
    out = hairsim.DeviceBuffer(16 * S * M)
    try:
        gpu.resample(M, out)
        ctx.sync()
        host, _ = gpu.read()
        want = oracle.resample32(host, N, M)
        assert_bitwise(out.numpy(np.float32), want.reshape(-1), f"resample S={S} N={N} M={M}")
    finally:
        out.free()
        gpu.destroy()


def test_resample_rejects_one_sample(ctx):
    wl = scenes.plain_1k(S=64)
    gpu = make_gpu_groom(ctx, wl)
    out = hairsim.DeviceBuffer(16 * 64)
    try:
        with pytest.raises(hairsim.HairSimError) as e:
            gpu.resample(1, out)
        assert e.value.code == hairsim.HS_ERR_INVALID
    finally:
        out.free()
        gpu.destroy()
