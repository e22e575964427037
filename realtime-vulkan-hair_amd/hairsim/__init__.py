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
LIB_PATH = os.environ.get("HAIRSIM_LIB") or os.path.join(_HERE, "libhairsim.so")

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
        ("forcing", C.c_uint32),
    ]


class HsMeshDesc(C.Structure):
    _fields_ = [("grid_dim", C.c_uint32 * 3), ("margin", C.c_float)]


class HsTile(C.Structure):
    _fields_ = [("gather_cells", C.c_uint32), ("splat_cells", C.c_uint32), ("xcd_order", C.c_uint32),
                ("strands", C.c_uint32), ("points", C.c_uint32)]


# (strands per workgroup, points per segment) of the built step kernels
TILE_SHAPES = ((128, 16), (256, 8), (256, 12))


class HsBatchItem(C.Structure):
    _fields_ = [
        ("groom_id", C.c_uint32),
        ("time", HsTime),
        ("colliders", C.POINTER(HsCollider)),
        ("num_colliders", C.c_uint32),
        ("params", C.POINTER(HsStepParams)),
    ]


class HsKeyframe(C.Structure):
    _fields_ = [("t", C.c_float), ("wind", C.c_float * 3), ("sphere", C.c_float * 3)]


class HsJob(C.Structure):
    _fields_ = [
        ("groom", HsGroomDesc),
        ("strands_aos", C.POINTER(C.c_float)),
        ("steps", C.c_uint32),
        ("dt", C.c_float),
        ("keys", C.POINTER(HsKeyframe)),
        ("num_keys", C.c_uint32),
        ("params", HsStepParams),
        ("result_aos", C.POINTER(C.c_float)),
    ]


class HsServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_uint64), ("max_points_per_step", C.c_uint64),
                ("use_graphs", C.c_int), ("step_loops", C.c_uint32), ("launch_depth", C.c_uint32), ("tile", HsTile)]


class HsServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64),
        ("launches", C.c_uint64), ("job_steps", C.c_uint64), ("point_updates", C.c_uint64),
        ("p50_step_ms", C.c_double), ("p99_step_ms", C.c_double), ("max_step_ms", C.c_double),
        ("p50_wait_ms", C.c_double), ("p99_wait_ms", C.c_double), ("busy_s", C.c_double),
        ("peak_resident_bytes", C.c_uint64), ("peak_jobs_per_launch", C.c_uint32),
        ("gpu_busy_s", C.c_double), ("p50_latency_ms", C.c_double), ("p99_latency_ms", C.c_double),
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
    "hs_read_strands_noalloc": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_float), C.POINTER(HsDrawIndirect)]),
    "hs_write_strands": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_float)]),
    "hs_read_grid": (C.c_int, [_P, C.c_uint32, C.POINTER(HsGridCell)]),
    "hs_grid_near_limit": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "hs_grid_work": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32 * 6)]),
    "hs_set_batch_graphs": (C.c_int, [_P, C.c_int]),
    "hs_batch_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "hs_set_tile": (C.c_int, [_P, C.POINTER(HsTile)]),
    "hs_get_tile": (C.c_int, [_P, C.POINTER(HsTile)]),
    "hs_autotune_batch": (C.c_int, [_P, C.POINTER(HsBatchItem), C.c_uint32, C.c_uint32, C.POINTER(HsTile),
                                    C.POINTER(C.c_double)]),
    "hs_mesh_create": (C.c_int, [_P, C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                 C.POINTER(HsMeshDesc), C.POINTER(C.c_uint32)]),
    "hs_mesh_destroy": (C.c_int, [_P, C.c_uint32]),
    "hs_profile_enable": (C.c_int, [_P, C.c_int]),
    "hs_profile_read": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "hs_profile_breakdown": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hs_profile_mesh": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "hs_export_strands": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_void_p]),
    "hs_obj_triangles": (C.c_int, [C.c_char_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                   C.POINTER(C.c_uint32)]),
    "hs_sample_roots": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "hs_initial_strands": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_float,
                                     C.POINTER(C.c_float)]),
    "hs_resample_strands": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p]),
    "hs_collider_from_trs": (None, [C.POINTER(HsCollider), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]),
    "hs_collider_translate": (None, [C.POINTER(HsCollider), C.POINTER(C.c_float)]),
    "hs_server_create": (C.c_int, [C.POINTER(HsServerConfig), C.POINTER(_P)]),
    "hs_server_submit": (C.c_int, [_P, C.POINTER(HsJob), C.POINTER(C.c_uint64)]),
    "hs_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "hs_server_cancel_pending": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "hs_server_drain": (C.c_int, [_P]),
    "hs_server_get_stats": (C.c_int, [_P, C.POINTER(HsServerStats)]),
    "hs_server_error": (C.c_int, [_P, C.POINTER(C.c_char_p)]),
    "hs_server_destroy": (None, [_P]),
    "hs_keyframes_eval": (None, [C.POINTER(HsKeyframe), C.c_uint32, C.c_float, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
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
                sphere_radius=1.0, vmax=10.0, friction=0.08, num_colliders=0, hair_hair=True, mesh_id=0,
                forcing=0):
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
    p.forcing = forcing
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

    def set_batch_graphs(self, on):
        self._check(lib().hs_set_batch_graphs(self._h, 1 if on else 0))

    def batch_stats(self):
        a = C.c_uint64()
        b = C.c_uint64()
        self._check(lib().hs_batch_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def batch(self, grooms, max_colliders=1):
        return Batch(self, grooms, max_colliders)

    def set_tile(self, gather_cells=2048, splat_cells=2048, xcd_order=1, strands=128, points=16):
        t = HsTile(gather_cells, splat_cells, xcd_order, strands, points)
        self._check(lib().hs_set_tile(self._h, C.byref(t)))

    def get_tile(self):
        """(gather_cells, splat_cells, xcd_order, strands, points)"""
        t = HsTile()
        self._check(lib().hs_get_tile(self._h, C.byref(t)))
        return (t.gather_cells, t.splat_cells, t.xcd_order, t.strands, t.points)

    def autotune(self, batch, steps=3):
        """Time every tile candidate on copies of `batch`'s grooms; keep the fastest.
        Returns ((gather_cells, splat_cells, xcd_order, strands, points), ms per batched step)."""
        t = HsTile()
        ms = C.c_double()
        self._check(lib().hs_autotune_batch(self._h, batch._items, batch.n, steps, C.byref(t), C.byref(ms)))
        return (t.gather_cells, t.splat_cells, t.xcd_order, t.strands, t.points), ms.value

    def mesh(self, verts, tris, grid_dim=(16, 16, 16), margin=0.05):
        """Register a closed triangle mesh collider; returns its id for step_params(mesh_id=...)."""
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        d = HsMeshDesc((C.c_uint32 * 3)(*grid_dim), margin)
        mid = C.c_uint32()
        self._check(lib().hs_mesh_create(self._h, _fptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         t.shape[0], C.byref(d), C.byref(mid)))
        return mid.value

    def mesh_destroy(self, mesh_id):
        self._check(lib().hs_mesh_destroy(self._h, mesh_id))

    def profile_breakdown(self):
        a = C.c_double()
        b = C.c_double()
        self._check(lib().hs_profile_breakdown(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def profile_mesh(self):
        """Summed ms of the mesh pass (hs_step with a mesh collider)."""
        a = C.c_double()
        self._check(lib().hs_profile_mesh(self._h, C.byref(a)))
        return a.value

    def groom(self, strands_aos, num_points, *, strand_length=2.5, seg_length=None, wind_noise=None,
              noise_dir=(1.0, 0.0, 0.0), grid_dim=0, grid_origin=(0.0, 0.0, 0.0), grid_h=1.0,
              keep_corrections=False):
        return Groom(self, strands_aos, num_points, strand_length=strand_length, seg_length=seg_length,
                     wind_noise=wind_noise, noise_dir=noise_dir, grid_dim=grid_dim,
                     grid_origin=grid_origin, grid_h=grid_h, keep_corrections=keep_corrections)


def _struct_dtype(st):
    """numpy view type of a ctypes Structure (pointers as uint64, nested structs flattened)."""
    names, formats, offsets = [], [], []
    for name, ct in st._fields_:
        off = getattr(st, name).offset
        if isinstance(ct, type) and issubclass(ct, C.Structure):
            names.append(name)
            formats.append(_struct_dtype(ct))
        elif isinstance(ct, type) and issubclass(ct, C.Array):
            names.append(name)
            base = {C.c_float: np.float32, C.c_uint32: np.uint32, C.c_int32: np.int32}[ct._type_]
            formats.append((base, (ct._length_,)))
        elif isinstance(ct, type) and issubclass(ct, C._Pointer):
            names.append(name)
            formats.append(np.uint64)
        else:
            names.append(name)
            formats.append({C.c_float: np.float32, C.c_uint32: np.uint32, C.c_int32: np.int32,
                            C.c_int: np.int32, C.c_uint64: np.uint64, C.c_double: np.float64}[ct])
        offsets.append(off)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(st)})


class Batch:
    """A fixed set of grooms stepped together by ``hs_step_batch``.

    The item, collider and parameter arrays are allocated once; numpy views
    (``time``, ``colliders``, ``wind``, ``num_colliders``, ``params``) update
    them in place between steps, so a step costs one C call.
    """

    def __init__(self, ctx, grooms, max_colliders=1):
        n = len(grooms)
        self.ctx = ctx
        self.grooms = list(grooms)
        self.n = n
        self.max_colliders = max_colliders
        self._items = (HsBatchItem * n)()
        self._cols = (HsCollider * (n * max(1, max_colliders)))()
        self._params = (HsStepParams * n)()
        base = C.addressof(self._cols)
        for k, g in enumerate(grooms):
            self._params[k] = step_params()
            it = self._items[k]
            it.groom_id = g.id
            it.colliders = C.cast(base + k * max(1, max_colliders) * C.sizeof(HsCollider), C.POINTER(HsCollider))
            it.num_colliders = 0
            it.params = C.pointer(self._params[k])
        self.items = np.frombuffer(self._items, dtype=_struct_dtype(HsBatchItem))
        self.params = np.frombuffer(self._params, dtype=_struct_dtype(HsStepParams))
        self.colliders = np.frombuffer(self._cols, dtype=np.float32).reshape(n, max(1, max_colliders), 48)

    def set_time(self, dt, total_time):
        self.items["time"]["deltaTime"] = dt
        self.items["time"]["totalTime"] = total_time

    def set_colliders(self, cols, num=None):
        """cols: (n, m, 48) float32; every groom gets m colliders (params.num_colliders too)."""
        cols = np.asarray(cols, dtype=np.float32)
        m = cols.shape[1]
        assert m <= self.max_colliders
        self.colliders[:, :m] = cols
        self.items["num_colliders"] = m
        self.params["num_colliders"] = m if num is None else num

    def step(self):
        self.ctx._check(lib().hs_step_batch(self.ctx._h, self._items, self.n))


def obj_triangles(path):
    """(T, 3, 3) corner positions and (T, 3, 3) corner normals of an OBJ file
    (triangulated as tinyobjloader 2.0 does)."""
    n = C.c_uint32()
    rc = lib().hs_obj_triangles(os.fsencode(path), None, None, 0, C.byref(n))
    if rc != HS_OK:
        raise HairSimError(rc, f"hs_obj_triangles({path!r})")
    pos = np.empty((n.value, 3, 3), dtype=np.float32)
    nrm = np.empty((n.value, 3, 3), dtype=np.float32)
    rc = lib().hs_obj_triangles(os.fsencode(path), _fptr(pos), _fptr(nrm), n.value, C.byref(n))
    if rc != HS_OK:
        raise HairSimError(rc, f"hs_obj_triangles({path!r})")
    return pos, nrm


def sample_roots(positions, normals, num_strands, seed=8):
    """Roots and root normals (S, 3) as GeneratePointsOnMesh samples them."""
    p = np.ascontiguousarray(positions, dtype=np.float32)
    n = np.ascontiguousarray(normals, dtype=np.float32)
    T = p.size // 9
    roots = np.empty((num_strands, 3), dtype=np.float32)
    rn = np.empty((num_strands, 3), dtype=np.float32)
    rc = lib().hs_sample_roots(_fptr(p), _fptr(n), T, num_strands, seed, _fptr(roots), _fptr(rn))
    if rc != HS_OK:
        raise HairSimError(rc, "hs_sample_roots")
    return roots, rn


def initial_strands(roots, root_normals, num_points, length=2.5):
    """Strand-layout buffer of Hair::Hair's initial strands."""
    r = np.ascontiguousarray(roots, dtype=np.float32)
    n = np.ascontiguousarray(root_normals, dtype=np.float32)
    S = r.size // 3
    out = np.empty(S * 12 * num_points, dtype=np.float32)
    rc = lib().hs_initial_strands(_fptr(r), _fptr(n), S, num_points, length, _fptr(out))
    if rc != HS_OK:
        raise HairSimError(rc, "hs_initial_strands")
    return out


_hip = None


def _hip_runtime():
    """The HIP runtime libhairsim.so is linked against (already in the process)."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip = H
    return _hip


class DeviceBuffer:
    """Device memory for the device-output entry points (hs_export_strands,
    hs_resample_strands), allocated with the library's own HIP runtime."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        rc = _hip_runtime().hipMalloc(C.byref(p), max(1, self.nbytes))
        if rc != 0:
            raise HairSimError(HS_ERR_OUT_OF_MEMORY, f"hipMalloc({self.nbytes}) failed: {rc}")
        self.ptr = p.value

    def fill(self, byte):
        _hip_runtime().hipMemset(self.ptr, int(byte), self.nbytes)

    def numpy(self, dtype):
        """Blocking copy to the host.  It does not wait for the library's
        stream: call Context.sync() after the calls that write the buffer."""
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        rc = _hip_runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2)  # device to host
        if rc != 0:
            raise HairSimError(HS_ERR_HIP, f"hipMemcpy failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            _hip_runtime().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
        return False


def _dev(buf):
    if buf is None:
        return None
    return buf.ptr if isinstance(buf, DeviceBuffer) else int(buf)


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

    def export(self, strands_dev, indirect_dev=None):
        """Reference-layout strands (and the indirect record) into device memory
        (DeviceBuffer or a raw device address); enqueued, ctx.sync() waits."""
        self.ctx._check(lib().hs_export_strands(self.ctx._h, self.id, _dev(strands_dev), _dev(indirect_dev)))

    def resample(self, samples, out_dev):
        """Tessellation-shader curve samples (S x samples float4) into device memory."""
        self.ctx._check(lib().hs_resample_strands(self.ctx._h, self.id, int(samples), _dev(out_dev)))

    def write(self, strands_aos):
        aos = np.ascontiguousarray(strands_aos, dtype=np.float32)
        assert aos.size == self.S * 12 * self.N
        self.ctx._check(lib().hs_write_strands(self.ctx._h, self.id, _fptr(aos)))

    def read_grid(self):
        out = np.zeros(self.G ** 3 * 4, dtype=np.int32)
        self.ctx._check(lib().hs_read_grid(self.ctx._h, self.id,
                                           out.ctypes.data_as(C.POINTER(HsGridCell))))
        return out

    def grid_work(self):
        """Last splatting step: dict of pool cells / capacity, boxes / capacity, fallback flag,
        deferred (multi-window) workgroup-segments."""
        v = (C.c_uint32 * 6)()
        self.ctx._check(lib().hs_grid_work(self.ctx._h, self.id, C.byref(v)))
        return dict(zip(["pool_cells", "pool_capacity", "boxes", "box_capacity", "fallback", "deferred"], list(v)))

    def grid_near_limit(self):
        """(cell, step) pairs so far with a GridCell component within 2x of the int32 limit."""
        v = C.c_uint64(0)
        self.ctx._check(lib().hs_grid_near_limit(self.ctx._h, self.id, C.byref(v)))
        return v.value

    def destroy(self):
        if self.id:
            self.ctx._check(lib().hs_groom_destroy(self.ctx._h, self.id))
            self.id = 0


# --------------------------------------------------------------------------
# Served jobs
# --------------------------------------------------------------------------
def keyframes(keys):
    """(K, 7) array-like of [t, wind xyz, sphere xyz] -> ctypes HsKeyframe array."""
    k = np.ascontiguousarray(keys, dtype=np.float32).reshape(-1, 7)
    arr = (HsKeyframe * len(k))()
    np.frombuffer(arr, dtype=np.float32).reshape(-1, 7)[:] = k
    return arr


def keyframes_eval(keys, t):
    arr = keys if isinstance(keys, C.Array) else keyframes(keys)
    w = (C.c_float * 3)()
    s = (C.c_float * 3)()
    lib().hs_keyframes_eval(arr, len(arr), C.c_float(t), w, s)
    return tuple(w), tuple(s)


class Server:
    """Continuous-batching job server (hs_server_*): submit jobs, wait for their results."""

    def __init__(self, device=0, hbm_budget=1 << 30, max_points_per_step=1 << 24, use_graphs=False, step_loops=0,
                 launch_depth=0, tile=None):
        """tile: (gather_cells, splat_cells, xcd_order, strands, points) of every step loop, or None for the
        server's default (256 x 8)."""
        cfg = HsServerConfig(device, hbm_budget, max_points_per_step, 1 if use_graphs else 0, step_loops, launch_depth,
                             HsTile(*tile) if tile is not None else HsTile())
        self._h = C.c_void_p()
        rc = lib().hs_server_create(C.byref(cfg), C.byref(self._h))
        if rc != HS_OK:
            raise HairSimError(rc, "hs_server_create failed")
        self._keep = {}

    def _check(self, rc, what):
        if rc != HS_OK:
            msg = C.c_char_p()
            lib().hs_server_error(self._h, C.byref(msg))
            raise HairSimError(rc, f"{what}: {(msg.value or b'').decode()}")

    def submit(self, strands_aos, num_points, steps, dt, keys, params, *, strand_length=2.5, seg_length=None,
               wind_noise=None, noise_dir=(1.0, 0.0, 0.0), grid_dim=0, grid_origin=(0.0, 0.0, 0.0), grid_h=1.0,
               keep_result=True):
        """Queue one job; returns (job_id, result array filled when the job is done, or None
        with keep_result=False)."""
        aos = np.ascontiguousarray(strands_aos, dtype=np.float32).reshape(-1)
        S = aos.size // (12 * num_points)
        desc = HsGroomDesc()
        desc.num_strands = S
        desc.num_points = num_points
        desc.strand_length = strand_length
        seg = None if seg_length is None else np.ascontiguousarray(seg_length, dtype=np.float32)
        noise = None if wind_noise is None else np.ascontiguousarray(wind_noise, dtype=np.float32)
        desc.seg_length = _fptr(seg) if seg is not None else None
        desc.wind_noise = _fptr(noise) if noise is not None else None
        desc.noise_dir = (C.c_float * 3)(*noise_dir)
        desc.grid_dim = grid_dim
        desc.grid_origin = (C.c_float * 3)(*grid_origin)
        desc.grid_h = grid_h
        desc.flags = 0
        karr = keys if isinstance(keys, C.Array) else keyframes(keys)
        result = np.empty_like(aos) if keep_result else None
        job = HsJob()
        job.groom = desc
        job.strands_aos = _fptr(aos)
        job.steps = steps
        job.dt = dt
        job.keys = C.cast(karr, C.POINTER(HsKeyframe))
        job.num_keys = len(karr)
        job.params = params
        job.result_aos = _fptr(result) if result is not None else None
        jid = C.c_uint64()
        self._check(lib().hs_server_submit(self._h, C.byref(job), C.byref(jid)), "hs_server_submit")
        self._keep[jid.value] = (aos, seg, noise, karr, result)
        return jid.value, result

    def wait(self, job_id, timeout=-1.0):
        rc = lib().hs_server_wait(self._h, job_id, float(timeout))
        if rc == HS_OK:
            return self._keep.pop(job_id)[4]
        self._check(rc, f"hs_server_wait({job_id})")

    def cancel_pending(self):
        n = C.c_uint64()
        self._check(lib().hs_server_cancel_pending(self._h, C.byref(n)), "hs_server_cancel_pending")
        return n.value

    def drain(self):
        self._check(lib().hs_server_drain(self._h), "hs_server_drain")

    def stats(self):
        s = HsServerStats()
        self._check(lib().hs_server_get_stats(self._h, C.byref(s)), "hs_server_get_stats")
        return {name: getattr(s, name) for name, _ in HsServerStats._fields_}

    def error(self):
        """(status, message) of the first failure on a server thread; (0, '') if none."""
        msg = C.c_char_p()
        rc = lib().hs_server_error(self._h, C.byref(msg))
        return rc, (msg.value or b"").decode()

    def close(self):
        if self._h:
            lib().hs_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
