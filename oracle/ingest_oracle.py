"""Groom ingestion of the reference demo, restated in Python / numpy binary32
(TEST INFRASTRUCTURE: used by tests/ only, never by the product).

Follows src/Strand.cpp:10-24 (random helpers), 26-146 (GeneratePointsOnMesh),
149-191 (Hair::Hair), and the parts of src/tiny_obj_loader.h (tinyobjloader
2.0, vendored by the reference) that decide the bits: tryParseDouble
(:805-929) + parseReal (:931-939) for numbers, exportGroupsToShape's
triangulation (:1329-1560) and pnpoly (:1314-1326) for polygons, fixIndex
(:739) for indices.
"""
import math

import numpy as np

f32 = np.float32


def parse_number(tok):
    """tryParseDouble then the cast to float (tiny_obj_loader.h:805-939)."""
    m = 0.0
    i = 0
    n = len(tok)
    neg = False
    lead_dot = False
    if i < n and tok[i] in "+-":
        neg = tok[i] == "-"
        i += 1
        lead_dot = i < n and tok[i] == "."
    elif i < n and tok[i] == ".":
        lead_dot = True
    elif not (i < n and tok[i].isdigit()):
        return f32(0.0)
    if not lead_dot:
        while i < n and tok[i].isdigit():
            m = m * 10
            m = m + float(ord(tok[i]) - 48)
            i += 1
    ex = 0
    if i < n and tok[i] == ".":
        i += 1
        k = 1
        lut = [1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001]
        while i < n and tok[i].isdigit():
            m = m + float(ord(tok[i]) - 48) * (lut[k] if k < 8 else math.pow(10.0, -k))
            k += 1
            i += 1
    if i < n and tok[i] in "eE":
        i += 1
        eneg = False
        if i < n and tok[i] in "+-":
            eneg = tok[i] == "-"
            i += 1
        while i < n and tok[i].isdigit():
            ex = ex * 10 + ord(tok[i]) - 48
            i += 1
        ex = -ex if eneg else ex
    v = (-1.0 if neg else 1.0) * (math.ldexp(m * math.pow(5.0, ex), ex) if ex else m)
    return f32(v)


def _pnpoly(vx, vy, tx, ty):
    """W. R. Franklin's crossing test (tiny_obj_loader.h:1314-1326), binary32."""
    c = False
    j = 2
    for i in range(3):
        if (vy[i] > ty) != (vy[j] > ty):
            if tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]:
                c = not c
        j = i
    return c


def triangulate(poly, V):
    """Ear clipping of one polygon (list of (v, vn) corners), as
    exportGroupsToShape with triangulate = true (tiny_obj_loader.h:1355-1560)."""
    n = len(poly)
    if n < 3:
        return []
    if n == 3:
        return list(poly)
    P = lambda c, a: V[c[0]][a]
    ax = [1, 2]
    eps = f32(np.finfo(np.float32).eps)
    for k in range(n):
        a, b, c = poly[k], poly[(k + 1) % n], poly[(k + 2) % n]
        e0 = [P(b, q) - P(a, q) for q in range(3)]
        e1 = [P(c, q) - P(b, q) for q in range(3)]
        cx = abs(e0[1] * e1[2] - e0[2] * e1[1])
        cy = abs(e0[2] * e1[0] - e0[0] * e1[2])
        cz = abs(e0[0] * e1[1] - e0[1] * e1[0])
        if cx > eps or cy > eps or cz > eps:
            if not (cx > cy and cx > cz):
                ax[0] = 0
                if cz > cx and cz > cy:
                    ax[1] = 1
            break
    area = f32(0.0)
    for k in range(n):
        a, b = poly[k], poly[(k + 1) % n]
        area = area + (P(a, ax[0]) * P(b, ax[1]) - P(a, ax[1]) * P(b, ax[0])) * f32(0.5)
    rem = list(poly)
    out = []
    guess, iters, prev = 0, n, n
    while len(rem) > 3 and iters > 0:
        np_ = len(rem)
        if guess >= np_:
            guess -= np_
        if prev != np_:
            prev, iters = np_, np_
        else:
            iters -= 1
        tri = [rem[(guess + k) % np_] for k in range(3)]
        vx = [P(t, ax[0]) for t in tri]
        vy = [P(t, ax[1]) for t in tri]
        cr = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1])
        if cr * area < 0:
            guess += 1
            continue
        if any(_pnpoly(vx, vy, P(rem[(guess + o) % np_], ax[0]), P(rem[(guess + o) % np_], ax[1]))
               for o in range(3, np_)):
            guess += 1
            continue
        out += tri
        del rem[(guess + 1) % np_]
    if len(rem) == 3:
        out += rem
    return out


def _fix(idx, n):
    """fixIndex (tiny_obj_loader.h:739): 1-based, negative from the end."""
    if idx > 0:
        return idx - 1
    if idx == 0:
        raise ValueError("index 0")
    return n + idx


def obj_triangles(text):
    """(T, 3, 3) corner positions and normals of OBJ text, file order."""
    V, VN, tris = [], [], []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            V.append([parse_number(x) for x in (t[1:4] + ["0"] * 3)[:3]])
        elif t[0] == "vn":
            VN.append([parse_number(x) for x in (t[1:4] + ["0"] * 3)[:3]])
        elif t[0] == "f":
            poly = []
            for tok in t[1:]:
                parts = tok.split("/")
                v = _fix(int(parts[0]), len(V))
                vn = _fix(int(parts[2]), len(VN)) if len(parts) == 3 and parts[2] else -1
                poly.append((v, vn))
            tris += triangulate(poly, V)
    T = len(tris) // 3
    pos = np.array([[V[c[0]] for c in tris[3 * i:3 * i + 3]] for i in range(T)], dtype=np.float32).reshape(T, 3, 3)
    nrm = np.array([[VN[c[1]] for c in tris[3 * i:3 * i + 3]] for i in range(T)], dtype=np.float32).reshape(T, 3, 3)
    return pos, nrm


class MsvcRand:
    """MSVC rand() after srand(seed): x = x*214013 + 2531011 mod 2^32, bits 16..30."""

    def __init__(self, seed):
        self.x = seed & 0xFFFFFFFF

    def __call__(self):
        self.x = (self.x * 214013 + 2531011) & 0xFFFFFFFF
        return (self.x >> 16) & 0x7FFF


def sample_roots(pos, nrm, S, seed=8):
    """GeneratePointsOnMesh's sampling loop (Strand.cpp:88-107)."""
    r = MsvcRand(seed)
    T = pos.shape[0]
    roots = np.zeros((S, 3), dtype=np.float32)
    rn = np.zeros((S, 3), dtype=np.float32)
    for i in range(S):
        t = r() % T
        u = f32(r()) / f32(32767.0)
        v = f32(r()) / f32(32767.0)
        if u + v >= f32(1.0):
            u = f32(1.0) - u
            v = f32(1.0) - v
        w = (f32(1.0) - u) - v
        roots[i] = (pos[t, 0] * u + pos[t, 1] * v) + pos[t, 2] * w
        rn[i] = nrm[t, 0]
    return roots, rn


def initial_strands(roots, rn, N, length=2.5):
    """Hair::Hair's strands (Strand.cpp:157-172), Strand layout."""
    S = roots.shape[0]
    out = np.zeros((S, 3, N, 4), dtype=np.float32)
    step = f32(float(f32(length)) / (N - 1.0))
    d = rn.astype(np.float32).copy()
    d[:, 2] = d[:, 2] - f32(2.0)
    d[:, 1] = d[:, 1] + f32(5.0)
    d[:, 0] = d[:, 0] + f32(0.05)
    cp = roots.astype(np.float32).copy()
    for j in range(N):
        out[:, 0, j, :3] = cp
        out[:, 0, j, 3] = 1.0
        out[:, 1, j, 2] = -1.0
        cp = cp + step * d
    return out.reshape(-1)
