"""ctypes binding of the CPU oracle (oracle/libhair_oracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the parity reference.  It restates
src/shaders/compute.comp:124-303 (see hair_oracle.h for the arithmetic
contract); ``step32`` is "ref32", ``step64`` is "oracle64".
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhair_oracle.so")


class HoParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("gravity", C.c_float * 3),
        ("wind", C.c_float * 3),
        ("damping", C.c_float),
        ("spring_k", C.c_float),
        ("sphere_radius", C.c_float),
        ("vmax", C.c_float),
        ("friction", C.c_float),
        ("hair_hair", C.c_int),
        ("grid_dim", C.c_int),
        ("grid_origin", C.c_float * 3),
        ("grid_h", C.c_float),
        ("num_colliders", C.c_int),
        ("colliders", C.POINTER(C.c_float)),
        ("seg_length", C.POINTER(C.c_float)),
        ("wind_noise", C.POINTER(C.c_float)),
        ("noise_dir", C.c_float * 3),
        ("mesh_tris", C.POINTER(C.c_float)),
        ("mesh_ntris", C.c_int),
        ("mesh_margin", C.c_float),
        ("total_time", C.c_float),
        ("forcing", C.c_int),
    ]

MESH_TRI_FLOATS = 40


_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        FP = C.POINTER(C.c_float)
        DP = C.POINTER(C.c_double)
        IP = C.POINTER(C.c_int32)
        PP = C.POINTER(HoParams)
        L.ho_step_f32.argtypes = [PP, FP, C.c_int, C.c_int, IP]
        L.ho_step_f64.argtypes = [PP, DP, C.c_int, C.c_int, IP]
        L.ho_step_f32_mt.argtypes = [PP, FP, C.c_int, C.c_int, IP, C.c_int]
        L.ho_splat_point_f32.argtypes = [PP, FP, FP, IP]
        L.ho_gather_point_f32.argtypes = [PP, FP, IP, FP]
        L.ho_splat_point_f64.argtypes = [PP, DP, DP, IP]
        L.ho_gather_point_f64.argtypes = [PP, DP, IP, DP]
        L.ho_force_f32.argtypes = [PP, FP, C.c_float, FP]
        L.ho_force_f32.restype = C.c_int
        L.ho_force_f64.argtypes = [PP, DP, C.c_double, DP]
        L.ho_force_f64.restype = C.c_int
        L.ho_mesh_prepare.argtypes = [FP, C.c_int, C.POINTER(C.c_uint32), C.c_int, FP]
        L.ho_mesh_prepare.restype = C.c_int
        L.ho_sat_i32_f32.argtypes = [C.c_float]
        L.ho_sat_i32_f32.restype = C.c_int32
        L.ho_sat_i32_f64.argtypes = [C.c_double]
        L.ho_sat_i32_f64.restype = C.c_int32
        for fn in ("ho_sinf", "ho_cosf"):
            getattr(L, fn).argtypes = [C.c_float]
            getattr(L, fn).restype = C.c_float
        L.ho_fbmf.argtypes = [C.c_float, C.c_float]
        L.ho_fbmf.restype = C.c_float
        L.ho_resample_f32.argtypes = [FP, C.c_int, C.c_int, C.c_int, FP]
        _lib = L
    return _lib


class Params:
    """Owns an HoParams and the arrays it points to."""

    def __init__(self, *, dt, gravity=(0.0, -9.8, 0.0), wind=(0.0, 0.0, 0.0), damping=0.998, spring_k=1900.0,
                 sphere_radius=1.0, vmax=10.0, friction=0.08, hair_hair=False, grid_dim=0,
                 grid_origin=(0.0, 0.0, 0.0), grid_h=1.0, colliders=None, seg_length=None, wind_noise=None,
                 noise_dir=(1.0, 0.0, 0.0), mesh=None, mesh_margin=0.0, total_time=0.0, forcing=0):
        p = HoParams()
        p.dt = dt
        p.gravity = (C.c_float * 3)(*gravity)
        p.wind = (C.c_float * 3)(*wind)
        p.damping = damping
        p.spring_k = spring_k
        p.sphere_radius = sphere_radius
        p.vmax = vmax
        p.friction = friction
        p.hair_hair = 1 if hair_hair else 0
        p.grid_dim = grid_dim
        p.grid_origin = (C.c_float * 3)(*grid_origin)
        p.grid_h = grid_h
        self._cols = np.ascontiguousarray(np.asarray(colliders if colliders is not None else np.zeros((0, 48)),
                                                     dtype=np.float32).reshape(-1, 48))
        p.num_colliders = self._cols.shape[0]
        p.colliders = self._cols.ctypes.data_as(C.POINTER(C.c_float))
        self._seg = None if seg_length is None else np.ascontiguousarray(seg_length, dtype=np.float32)
        p.seg_length = None if self._seg is None else self._seg.ctypes.data_as(C.POINTER(C.c_float))
        self._noise = None if wind_noise is None else np.ascontiguousarray(wind_noise, dtype=np.float32)
        p.wind_noise = None if self._noise is None else self._noise.ctypes.data_as(C.POINTER(C.c_float))
        p.noise_dir = (C.c_float * 3)(*noise_dir)
        # mesh: the (T, 40) records of mesh_prepare
        self._mesh = None if mesh is None else np.ascontiguousarray(mesh, dtype=np.float32)
        p.mesh_tris = None if self._mesh is None else self._mesh.ctypes.data_as(C.POINTER(C.c_float))
        p.mesh_ntris = 0 if self._mesh is None else self._mesh.shape[0]
        p.mesh_margin = mesh_margin
        p.total_time = total_time
        p.forcing = forcing
        self.p = p

    def grid_buffer(self):
        G = self.p.grid_dim
        return np.zeros(max(1, G ** 3) * 4, dtype=np.int32)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def step32(params, strands, N, grid=None):
    """One ref32 step in place on float32 AoS strands (S, 12N)."""
    assert strands.dtype == np.float32 and strands.flags.c_contiguous
    S = strands.size // (12 * N)
    if grid is None:
        grid = params.grid_buffer()
    lib().ho_step_f32(C.byref(params.p), strands.ctypes.data_as(C.POINTER(C.c_float)), S, N, _ip(grid))
    return grid


def step32_mt(params, strands, N, grid=None, threads=0):
    assert strands.dtype == np.float32 and strands.flags.c_contiguous
    S = strands.size // (12 * N)
    if grid is None:
        grid = params.grid_buffer()
    lib().ho_step_f32_mt(C.byref(params.p), strands.ctypes.data_as(C.POINTER(C.c_float)), S, N, _ip(grid),
                         threads)
    return grid


def step64(params, strands, N, grid=None):
    """One oracle64 step in place on float64 AoS strands (S, 12N)."""
    assert strands.dtype == np.float64 and strands.flags.c_contiguous
    S = strands.size // (12 * N)
    if grid is None:
        grid = params.grid_buffer()
    lib().ho_step_f64(C.byref(params.p), strands.ctypes.data_as(C.POINTER(C.c_double)), S, N, _ip(grid))
    return grid


def splat32(params, pos, vel, grid):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    FP = C.POINTER(C.c_float)
    lib().ho_splat_point_f32(C.byref(params.p), pos.ctypes.data_as(FP), vel.ctypes.data_as(FP), _ip(grid))


def gather32(params, pos, grid):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    FP = C.POINTER(C.c_float)
    lib().ho_gather_point_f32(C.byref(params.p), pos.ctypes.data_as(FP), _ip(grid), out.ctypes.data_as(FP))
    return out


def splat64(params, pos, vel, grid):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    DP = C.POINTER(C.c_double)
    lib().ho_splat_point_f64(C.byref(params.p), pos.ctypes.data_as(DP), vel.ctypes.data_as(DP), _ip(grid))


def gather64(params, pos, grid):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    DP = C.POINTER(C.c_double)
    lib().ho_gather_point_f64(C.byref(params.p), pos.ctypes.data_as(DP), _ip(grid), out.ctypes.data_as(DP))
    return out


def force32(params, pos, noise=0.0):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    FP = C.POINTER(C.c_float)
    hits = lib().ho_force_f32(C.byref(params.p), pos.ctypes.data_as(FP), noise, out.ctypes.data_as(FP))
    return hits, out


def force64(params, pos, noise=0.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    DP = C.POINTER(C.c_double)
    hits = lib().ho_force_f64(C.byref(params.p), pos.ctypes.data_as(DP), noise, out.ctypes.data_as(DP))
    return hits, out


def resample32(strands, N, M):
    """Tessellation-shader curve samples: (S, M, 4) float32."""
    a = np.ascontiguousarray(strands, dtype=np.float32).reshape(-1)
    S = a.size // (12 * N)
    out = np.zeros((S, M, 4), dtype=np.float32)
    FP = C.POINTER(C.c_float)
    lib().ho_resample_f32(a.ctypes.data_as(FP), S, N, M, out.ctypes.data_as(FP))
    return out


def sat_i32(f):
    return int(lib().ho_sat_i32_f32(f))


def mesh_prepare(verts, tris):
    """Per-triangle records (vertices + pseudo-normals) of the oracle's mesh collider."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    out = np.zeros((t.shape[0], MESH_TRI_FLOATS), dtype=np.float32)
    rc = lib().ho_mesh_prepare(v.ctypes.data_as(C.POINTER(C.c_float)), v.shape[0],
                               t.ctypes.data_as(C.POINTER(C.c_uint32)), t.shape[0],
                               out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise ValueError("mesh is not a closed 2-manifold (or has an index out of range)")
    return out
