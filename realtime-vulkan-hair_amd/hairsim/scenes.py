"""Deterministic synthetic grooms and per-step parameters for the five workloads.

Strands are produced in the reference ``Strand`` layout (src/Strand.h:11-16):
``vec4 curvePoints[N]; vec4 curveVels[N]; vec4 correctionVecs[N]`` per strand,
with w = 1 for positions and 0 for velocities/corrections as
src/Strand.cpp:159-172 initialises them.  Initial strands are straight along
the root normal.  Initial velocity is 0 (the reference starts at (0,0,-1),
Strand.cpp:165; the workloads start at rest).

Time convention: like Scene::UpdateTime before Renderer::Frame
(src/main.cpp:282-283, src/Scene.cpp:78-87), step k (0-based) sees
``totalTime = (k + 1) * dt``; wind and collider trajectories are evaluated at
that time.

Workloads (all fp32):
  plain-1k      1024 x 16, L = 1, dt 1/60, g (0,-9.8,0), damping 0.9, no wind,
                sphere r 0.3 at (0,0,1.2), no hair-hair, 240 steps.
  plain-100k    100,000 x 32, L ~ U[0.8,1.2], wind x = 2 sin(2 pi t / 1 s) plus a
                per-strand constant N(0, 0.1^2) along x, sphere r 0.3 orbiting a
                circle of radius 1.1 in the plane y = 0.5 with period 4 s, 64^3
                grid over [-2.2, 2.2]^3, friction 0.05, 240 steps.
  batched-256   256 grooms, S log-uniform in [512, 16384], N in {8,16,32,64},
                per-groom wind ~ N(0, 1.5^2) per axis, own sphere, 32^3 grid.
  large-1m      1,000,000 x 64, as plain-100k with a 128^3 grid (+ head mesh).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 0x5EED0000


def hemisphere_roots(rng, S, radius=1.0, scale=(1.0, 1.0, 1.0)):
    """Area-uniform roots on the upper unit hemisphere and their outward normals.

    y ~ U[0,1], phi ~ U[0, 2 pi): root = (sqrt(1-y^2) cos phi, y, sqrt(1-y^2) sin phi).
    With a non-unit `scale` the roots are mapped onto the scaled hemisphere and
    the normals are the scaled surface's normals.
    """
    y = rng.random(S)
    phi = rng.random(S) * (2.0 * math.pi)
    rxz = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    unit = np.stack([rxz * np.cos(phi), y, rxz * np.sin(phi)], axis=1)
    sc = np.asarray(scale, dtype=np.float64)
    roots = unit * sc * radius
    nrm = unit / sc
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return roots, nrm


def straight_strands(roots, normals, N, lengths):
    """AoS strands (S, 12N float32): p_i = root + i * (L / (N-1)) * n, v = 0, corr = 0."""
    S = roots.shape[0]
    seg = np.asarray(lengths, dtype=np.float32) / np.float32(N - 1)
    i = np.arange(N, dtype=np.float64)
    pts = roots[:, None, :] + i[None, :, None] * seg.astype(np.float64)[:, None, None] * normals[:, None, :]
    aos = np.zeros((S, 3, N, 4), dtype=np.float32)
    aos[:, 0, :, :3] = pts.astype(np.float32)
    aos[:, 0, :, 3] = 1.0
    return aos.reshape(S, 12 * N), seg


@dataclass
class Groom:
    strands: np.ndarray            # (S, 12N) float32, reference layout
    N: int
    seg: np.ndarray                # (S,) float32 segment rest length
    noise: np.ndarray | None       # (S,) float32 per-strand wind noise or None
    noise_dir: tuple = (1.0, 0.0, 0.0)
    grid_dim: int = 0
    grid_origin: tuple = (-2.2, -2.2, -2.2)
    grid_h: float = 4.4 / 64

    @property
    def S(self):
        return self.strands.shape[0]


@dataclass
class Workload:
    name: str
    dt: float
    steps: int
    damping: float
    friction: float
    hair_hair: bool
    sphere_radius: float
    grooms: list = field(default_factory=list)
    wind_amp: float = 0.0          # x wind amplitude (sinusoid)
    wind_period: float = 1.0
    wind_const: list = field(default_factory=list)  # per-groom constant wind vectors
    sphere_center: tuple = (0.0, 0.0, 1.2)
    sphere_orbit: float = 0.0      # orbit radius (0: fixed centre)
    sphere_orbit_y: float = 0.5
    sphere_period: float = 4.0
    sphere_centers: list = field(default_factory=list)  # per-groom fixed centres
    mesh: tuple = None             # (verts (V,3) f32, tris (T,3) u32) head mesh collider, or None
    mesh_margin: float = 0.05
    mesh_grid: tuple = (32, 32, 32)

    def time(self, k):
        """totalTime seen by step k (0-based)."""
        return np.float32((k + 1) * np.float32(self.dt))

    def wind(self, k, groom=0):
        t = float(self.time(k))
        w = [0.0, 0.0, 0.0]
        if self.wind_amp:
            w[0] = self.wind_amp * math.sin(2.0 * math.pi * t / self.wind_period)
        if self.wind_const:
            c = self.wind_const[groom]
            w = [w[0] + c[0], w[1] + c[1], w[2] + c[2]]
        return tuple(np.float32(x) for x in w)

    def sphere(self, k, groom=0):
        if self.sphere_centers:
            return tuple(self.sphere_centers[groom])
        if self.sphere_orbit:
            t = float(self.time(k))
            a = 2.0 * math.pi * t / self.sphere_period
            return (self.sphere_orbit * math.cos(a), self.sphere_orbit_y, self.sphere_orbit * math.sin(a))
        return tuple(self.sphere_center)


def plain_1k(seed=SEED_BASE + 0, S=1024, N=16):
    rng = np.random.default_rng(seed)
    roots, nrm = hemisphere_roots(rng, S)
    aos, seg = straight_strands(roots, nrm, N, np.full(S, 1.0))
    g = Groom(aos, N, seg, None, grid_dim=0)
    return Workload("plain-1k", dt=1.0 / 60.0, steps=240, damping=0.9, friction=0.0, hair_hair=False,
                    sphere_radius=0.3, grooms=[g], sphere_center=(0.0, 0.0, 1.2))


def plain_100k(seed=SEED_BASE + 1, S=100_000, N=32, G=64):
    rng = np.random.default_rng(seed)
    roots, nrm = hemisphere_roots(rng, S)
    L = rng.uniform(0.8, 1.2, S)
    aos, seg = straight_strands(roots, nrm, N, L)
    noise = (rng.standard_normal(S) * 0.1).astype(np.float32)
    g = Groom(aos, N, seg, noise, grid_dim=G, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / G)
    return Workload("plain-100k", dt=1.0 / 60.0, steps=240, damping=0.9, friction=0.05, hair_hair=True,
                    sphere_radius=0.3, grooms=[g], wind_amp=2.0, wind_period=1.0, sphere_orbit=1.1)


def head_mesh(subdiv=5, scale=(1.0, 1.2, 1.0)):
    """The synthetic head: an icosphere (subdivision 5: 10,242 vertices,
    20,480 triangles) scaled per axis."""
    return icosphere(subdiv, scale)


def mesh_head_scene(name, seed, S, N, G, subdiv, scale=(1.0, 1.2, 1.0), margin=0.05, mesh_grid=(176, 209, 176),
                    steps=120):
    """Roots on the scaled upper hemisphere (the head's scalp), strands along
    the surface normal; wind, noise, orbiting sphere and grid as plain-100k,
    plus the head mesh collider.  The mesh grid only sets the contact search's
    speed, not its result (at 1M x 64, 110 steps, the mesh pass measured 2.04 /
    1.77 / 1.58 / 1.55 / 1.52 / 1.53 / 1.58 ms at 72, 96, 128, 144, 176, 192,
    224 cells along x)."""
    rng = np.random.default_rng(seed)
    roots, nrm = hemisphere_roots(rng, S, scale=scale)
    L = rng.uniform(0.8, 1.2, S)
    aos, seg = straight_strands(roots, nrm, N, L)
    noise = (rng.standard_normal(S) * 0.1).astype(np.float32)
    g = Groom(aos, N, seg, noise, grid_dim=G, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / G)
    return Workload(name, dt=1.0 / 60.0, steps=steps, damping=0.9, friction=0.05, hair_hair=True,
                    sphere_radius=0.3, grooms=[g], wind_amp=2.0, wind_period=1.0, sphere_orbit=1.1,
                    mesh=icosphere(subdiv, scale), mesh_margin=margin, mesh_grid=mesh_grid)


def large_1m(seed=SEED_BASE + 4, S=1_000_000, N=64, G=128):
    """1,000,000 strands x 64 points, 128^3 grid, head mesh of 20,480 triangles."""
    return mesh_head_scene("large-1m", seed, S, N, G, subdiv=5)


def batched(seed=SEED_BASE + 2, n_grooms=256, s_lo=512, s_hi=16384, n_choices=(8, 16, 32, 64), G=32):
    rng = np.random.default_rng(seed)
    grooms, winds, centers = [], [], []
    for _ in range(n_grooms):
        S = int(round(math.exp(rng.uniform(math.log(s_lo), math.log(s_hi)))))
        N = int(rng.choice(n_choices))
        roots, nrm = hemisphere_roots(rng, S)
        L = rng.uniform(0.8, 1.2, S)
        aos, seg = straight_strands(roots, nrm, N, L)
        grooms.append(Groom(aos, N, seg, None, grid_dim=G, grid_origin=(-2.2, -2.2, -2.2), grid_h=4.4 / G))
        winds.append(tuple(float(x) for x in rng.normal(0.0, 1.5, 3)))
        a = rng.uniform(0, 2 * math.pi)
        centers.append((1.1 * math.cos(a), float(rng.uniform(0.0, 1.0)), 1.1 * math.sin(a)))
    return Workload("batched-256", dt=1.0 / 60.0, steps=240, damping=0.9, friction=0.05, hair_hair=True,
                    sphere_radius=0.3, grooms=grooms, wind_const=winds, sphere_centers=centers)


def icosphere(subdiv, scale=(1.0, 1.0, 1.0)):
    """Unit icosphere with `subdiv` midpoint subdivisions (20 * 4^subdiv
    triangles, outward winding), scaled per axis.  Returns (verts (V,3) f32,
    tris (T,3) u32)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = (np.array(verts) * np.asarray(scale, dtype=np.float64)).astype(np.float32)
    return V, np.array(faces, dtype=np.uint32)


def sphere_collider(center):
    """48 floats of a reference `Collider` (Scene.h:23-39) that is a pure
    translation to `center`: transform = translate(c), inv = translate(-c),
    invTrans = transpose(inv).  Only transform[3].xyz is read for collider 0
    (compute.comp:162)."""
    c = np.asarray(center, dtype=np.float32)
    m = np.zeros(48, dtype=np.float32)
    for b in range(3):
        m[b * 16:(b + 1) * 16] = np.eye(4, dtype=np.float32).reshape(-1)
    m[12:15] = c
    m[16 + 12:16 + 15] = -c
    m[32 + 3], m[32 + 7], m[32 + 11] = -c[0], -c[1], -c[2]
    return m


def strand_bytes(S, N):
    """The strand-state part of a step's bytes: 48 per non-root point (pos+vel
    read and written), 16 per strand (root + rest length)."""
    return S * (48 * (N - 1) + 16)


def algorithmic_bytes(S, N, G, hair_hair):
    """Bytes one step must move: 48 per non-root point (pos+vel read and
    written), 16 per strand (root + rest length), 32 per voxel (clear + one
    read) when the grid is on."""
    b = S * (48 * (N - 1) + 16)
    if hair_hair and G:
        b += 32 * G ** 3
    return b
