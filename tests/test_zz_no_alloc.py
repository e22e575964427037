"""hs_read_strands_noalloc (the read-back that allocates nothing) against
hs_read_strands: two calls in a row, the second on a smaller groom, give the
same bits and the same indirect record; a groom larger than the fixed staging
array is refused with HS_ERR_INVALID.  Through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import hairsim
from hairsim import scenes
from helpers import assert_bitwise, gpu_params, make_gpu_groom

pytestmark = pytest.mark.gpu


def _read_noalloc(ctx, gr):
    out = np.full(gr.S * 12 * gr.N, np.nan, dtype=np.float32)
    ind = hairsim.HsDrawIndirect()
    rc = hairsim.lib().hs_read_strands_noalloc(ctx._h, gr.id, out.ctypes.data_as(C.POINTER(C.c_float)),
                                               C.byref(ind))
    assert rc == hairsim.HS_OK, hairsim.lib().hs_last_error(ctx._h)
    return out, (ind.vertexCount, ind.instanceCount, ind.firstVertex, ind.firstInstance)


@pytest.mark.parametrize("shape", [((20000, 32, 64), (900, 10, 16))])
def test_noalloc_read_equals_read(ctx, shape):
    for S, N, G in shape:  # the second groom is smaller than the first
        print(f"noalloc read: {S} x {N}, grid {G}", flush=True)
        wl = scenes.plain_100k(S=S, N=N, G=G)
        gr = make_gpu_groom(ctx, wl)
        try:
            for k in range(3):
                gr.step(wl.dt, float(wl.time(k)), [scenes.sphere_collider(wl.sphere(k))], gpu_params(wl, k))
            ref, ind_ref = gr.read()
            got, ind = _read_noalloc(ctx, gr)
            assert ind == ind_ref == (S, 1, 0, 0)
            assert_bitwise(got, ref, f"noalloc read {S} x {N}", N)
        finally:
            gr.destroy()


def test_noalloc_read_refuses_oversize(ctx):
    S, N = 44000, 32  # 44000 * 32 * 48 bytes > 64 MiB
    assert S * N * 48 > (64 << 20)
    wl = scenes.plain_100k(S=S, N=N, G=16)
    gr = make_gpu_groom(ctx, wl)
    try:
        out = np.zeros(S * 12 * N, dtype=np.float32)
        rc = hairsim.lib().hs_read_strands_noalloc(ctx._h, gr.id, out.ctypes.data_as(C.POINTER(C.c_float)), None)
        assert rc == hairsim.HS_ERR_INVALID
        assert b"staging" in hairsim.lib().hs_last_error(ctx._h)
        # the indirect record alone needs no staging
        ind = hairsim.HsDrawIndirect()
        assert hairsim.lib().hs_read_strands_noalloc(ctx._h, gr.id, None, C.byref(ind)) == hairsim.HS_OK
        assert ind.vertexCount == S
    finally:
        gr.destroy()
