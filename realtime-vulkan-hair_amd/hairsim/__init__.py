"""hairsim - Python binding of the MI355X strand-physics C ABI (include/hairsim.h).

This is a thin ctypes layer over ``libhairsim.so``: every call goes to the HIP
library, and there is no CPU fallback.  If the library is missing, importing
the binding raises :class:`LibraryNotBuilt`.

The object model mirrors the reference host code that drives the compute
pass:
  * :class:`Context` - the device/queue bring-up (src/Instance.cpp:315) plus the
    per-frame submit (src/Renderer.cpp:2309-2319);
  * :class:`Groom` - one ``Hair`` strand buffer (src/Strand.h:61-80);
  * :func:`collider_from_trs` / :func:`collider_translate` - ``Collider`` and
    ``Scene::translateSphere`` (src/Scene.h:23-39, src/Scene.cpp:110-136).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhairsim.so")

HS_OK = 0
HS_ERR_INVALID = -1
HS_ERR_NO_DEVICE = -2
HS_ERR_OUT_OF_MEMORY = -3
HS_ERR_HIP = -4
HS_ERR_NOT_FOUND = -5
HS_ERR_STATE = -6
HS_ERR_BUDGET = -7

HS_GROOM_KEEP_CORRECTIONS = 0x1
HS_MAX_COLLIDERS = 8
HS_MAX_POINTS = 256


class LibraryNotBuilt(RuntimeError):
    pass


class HairSimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hairsim error {code}: {msg}")
        self.code = code


# --------------------------------------------------------------------------
# ctypes mirrors of the boundary structs
# --------------------------------------------------------------------------
class HsTime(C.Structure):
    _fields_ = [("deltaTime", C.c_float), ("totalTime", C.c_float)]


class HsCollider(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("inv", C.c_float * 16), ("invTrans", C.c_float * 16)]

    def as_array(self):
        return np.array(list(self.transform) + list(self.inv) + list(self.invTrans), dtype=np.float32)


class HsGridCell(C.Structure):
    _fields_ = [("velocity", C.c_int32 * 3), ("density", C.c_int32)]


class HsDrawIndirect(C.Structure):
    _fields_ = [("vertexCount", C.c_uint32), ("instanceCount", C.c_uint32),
                ("firstVertex", C.c_uint32), ("firstInstance", C.c_uint32)]


class HsGroomDesc(C.Structure):
    _fields_ = [
        ("num_strands", C.c_uint32),
        ("num_points", C.c_uint32),
        ("strand_length", C.c_float),
        ("seg_length", C.POINTER(C.c_float)),
        ("wind_noise", C.POINTER(C.c_float)),
        ("noise_dir", C.c_float * 3),
        ("grid_dim", C.c_uint32),
        ("grid_origin", C.c_float * 3),
        ("grid_h", C.c_float),
        ("flags", C.c_uint32),
    ]


class HsStepParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3),
        ("wind", C.c_float * 3),
        ("damping", C.c_float),
        ("spring_k", C.c_float),
        ("sphere_radius", C.c_float),
        ("vmax", C.c_float),
        ("friction", C.c_float),
        ("num_colliders", C.c_uint32),
        ("hair_hair", C.c_uint32),
        ("mesh_id", C.c_uint32),
    ]


class HsBatchItem(C.Structure):
    _fields_ = [
        ("groom_id", C.c_uint32),
        ("time", HsTime),
        ("colliders", C.POINTER(HsCollider)),
        ("num_colliders", C.c_uint32),
        ("params", C.POINTER(HsStepParams)),
    ]


# Every symbol include/hairsim.h declares, with its ctypes signature.
_P = C.c_void_p
_SIGS = {
    "hs_step_params_reference": (None, [C.POINTER(HsStepParams)]),
    "hs_groom_desc_reference": (None, [C.POINTER(HsGroomDesc), C.c_uint32, C.c_uint32]),
    "hs_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "hs_destroy": (None, [_P]),
    "hs_last_error": (C.c_char_p, [_P]),
    "hs_sync": (C.c_int, [_P]),
    "hs_set_stream": (C.c_int, [_P, _P]),
    "hs_memory_used": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "hs_groom_create": (C.c_int, [_P, C.POINTER(HsGroomDesc), C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
    "hs_groom_destroy": (C.c_int, [_P, C.c_uint32]),
    "hs_groom_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "hs_step": (C.c_int, [_P, C.c_uint32, C.POINTER(HsTime), C.POINTER(HsCollider), C.c_uint32,
                          C.POINTER(HsStepParams)]),
    "hs_step_batch": (C.c_int, [_P, C.POINTER(HsBatchItem), C.c_uint32]),
    "hs_read_strands": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_float), C.POINTER(HsDrawIndirect)]),
    "hs_write_strands": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_float)]),
    "hs_read_grid": (C.c_int, [_P, C.c_uint32, C.POINTER(HsGridCell)]),
    "hs_profile_enable": (C.c_int, [_P, C.c_int]),
    "hs_profile_read": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "hs_profile_breakdown": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hs_collider_from_trs": (None, [C.POINTER(HsCollider), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]),
    "hs_collider_translate": (None, [C.POINTER(HsCollider), C.POINTER(C.c_float)]),
}

_lib = None


def lib():
    """Load libhairsim.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return sorted(_SIGS)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


# --------------------------------------------------------------------------
# Parameters
# --------------------------------------------------------------------------
def reference_step_params():
    p = HsStepParams()
    lib().hs_step_params_reference(C.byref(p))
    return p


def step_params(gravity=(0.0, -9.8, 0.0), wind=(0.0, 0.0, 0.0), damping=0.998, spring_k=1900.0,
                sphere_radius=1.0, vmax=10.0, friction=0.08, num_colliders=0, hair_hair=True, mesh_id=0):
    p = HsStepParams()
    p.gravity = _f3(gravity)
    p.wind = _f3(wind)
    p.damping = damping
    p.spring_k = spring_k
    p.sphere_radius = sphere_radius
    p.vmax = vmax
    p.friction = friction
    p.num_colliders = num_colliders
    p.hair_hair = 1 if hair_hair else 0
    p.mesh_id = mesh_id
    return p


def collider_from_trs(translation, rotation_deg=(0, 0, 0), scale=(1, 1, 1)):
    c = HsCollider()
    lib().hs_collider_from_trs(C.byref(c), _f3(translation), _f3(rotation_deg), _f3(scale))
    return c


def collider_translate(c, translation):
    lib().hs_collider_translate(C.byref(c), _f3(translation))
    return c


def groom_bytes(S, N, G=0, flags=0):
    return int(lib().hs_groom_bytes(S, N, G, flags))


# --------------------------------------------------------------------------
# Context / groom
# --------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().hs_create(C.byref(self._h), device)
        if rc != HS_OK:
            raise HairSimError(rc, "hs_create failed (no HIP device?)")

    def _check(self, rc):
        if rc != HS_OK:
            raise HairSimError(rc, lib().hs_last_error(self._h).decode())
        return rc

    def close(self):
        if self._h:
            lib().hs_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._check(lib().hs_sync(self._h))

    def memory_used(self):
        b = C.c_uint64()
        self._check(lib().hs_memory_used(self._h, C.byref(b)))
        return b.value

    def profile_enable(self, on=True):
        self._check(lib().hs_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        n = C.c_uint64()
        ms = C.c_double()
        self._check(lib().hs_profile_read(self._h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def step_batch(self, entries):
        """entries: list of (groom, dt, total_time, colliders, params)."""
        keep = []
        items = (HsBatchItem * max(1, len(entries)))()
        for k, (groom, dt, tt, cols, params) in enumerate(entries):
            cl = [c if isinstance(c, HsCollider) else HsCollider.from_buffer_copy(
                np.ascontiguousarray(c, dtype=np.float32).tobytes()) for c in (cols or [])]
            arr = (HsCollider * max(1, len(cl)))(*cl)
            keep.append((arr, params))
            items[k].groom_id = groom.id
            items[k].time = HsTime(dt, tt)
            items[k].colliders = C.cast(arr, C.POINTER(HsCollider))
            items[k].num_colliders = len(cl)
            items[k].params = C.pointer(params)
        self._check(lib().hs_step_batch(self._h, items, len(entries)))

    def profile_breakdown(self):
        a = C.c_double()
        b = C.c_double()
        self._check(lib().hs_profile_breakdown(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def groom(self, strands_aos, num_points, *, strand_length=2.5, seg_length=None, wind_noise=None,
              noise_dir=(1.0, 0.0, 0.0), grid_dim=0, grid_origin=(0.0, 0.0, 0.0), grid_h=1.0,
              keep_corrections=False):
        return Groom(self, strands_aos, num_points, strand_length=strand_length, seg_length=seg_length,
                     wind_noise=wind_noise, noise_dir=noise_dir, grid_dim=grid_dim,
                     grid_origin=grid_origin, grid_h=grid_h, keep_corrections=keep_corrections)


class Groom:
    def __init__(self, ctx, strands_aos, num_points, *, strand_length=2.5, seg_length=None, wind_noise=None,
                 noise_dir=(1.0, 0.0, 0.0), grid_dim=0, grid_origin=(0.0, 0.0, 0.0), grid_h=1.0,
                 keep_corrections=False):
        aos = np.ascontiguousarray(strands_aos, dtype=np.float32)
        N = int(num_points)
        S = aos.size // (12 * N)
        if aos.size != S * 12 * N:
            raise ValueError("strands_aos size is not a multiple of 12*N floats")
        self.ctx, self.S, self.N, self.G = ctx, S, N, int(grid_dim)
        d = HsGroomDesc()
        d.num_strands = S
        d.num_points = N
        d.strand_length = strand_length
        self._seg = self._noise = None
        if seg_length is not None:
            self._seg = np.ascontiguousarray(seg_length, dtype=np.float32)
            assert self._seg.size == S
            d.seg_length = _fptr(self._seg)
        if wind_noise is not None:
            self._noise = np.ascontiguousarray(wind_noise, dtype=np.float32)
            assert self._noise.size == S
            d.wind_noise = _fptr(self._noise)
        d.noise_dir = _f3(noise_dir)
        d.grid_dim = grid_dim
        d.grid_origin = _f3(grid_origin)
        d.grid_h = grid_h
        d.flags = HS_GROOM_KEEP_CORRECTIONS if keep_corrections else 0
        gid = C.c_uint32()
        ctx._check(lib().hs_groom_create(ctx._h, C.byref(d), _fptr(aos), C.byref(gid)))
        self.id = gid.value

    def step(self, dt, total_time, colliders, params):
        t = HsTime(dt, total_time)
        cols = [c if isinstance(c, HsCollider) else HsCollider.from_buffer_copy(
            np.ascontiguousarray(c, dtype=np.float32).tobytes()) for c in (colliders or [])]
        arr = (HsCollider * max(1, len(cols)))(*cols) if cols else None
        self.ctx._check(lib().hs_step(self.ctx._h, self.id, C.byref(t), arr, len(cols), C.byref(params)))

    def read(self):
        out = np.empty(self.S * 12 * self.N, dtype=np.float32)
        ind = HsDrawIndirect()
        self.ctx._check(lib().hs_read_strands(self.ctx._h, self.id, _fptr(out), C.byref(ind)))
        return out, (ind.vertexCount, ind.instanceCount, ind.firstVertex, ind.firstInstance)

    def write(self, strands_aos):
        aos = np.ascontiguousarray(strands_aos, dtype=np.float32)
        assert aos.size == self.S * 12 * self.N
        self.ctx._check(lib().hs_write_strands(self.ctx._h, self.id, _fptr(aos)))

    def read_grid(self):
        out = np.zeros(self.G ** 3 * 4, dtype=np.int32)
        self.ctx._check(lib().hs_read_grid(self.ctx._h, self.id,
                                           out.ctypes.data_as(C.POINTER(HsGridCell))))
        return out

    def destroy(self):
        if self.id:
            self.ctx._check(lib().hs_groom_destroy(self.ctx._h, self.id))
            self.id = 0
