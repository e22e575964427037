"""Groom ingestion (Strand.cpp:10-191 with tinyobjloader 2.0's parsing and
triangulation): the library's C++ against the Python restatement in
oracle/ingest_oracle.py, bit for bit, and the restatement pinned by known answers
and geometry."""
import math
import os

import numpy as np
import pytest

import hairsim
import ingest_oracle as ingest

MANNEQUIN = "/root/reference/src/models/mannequin_segment.obj"


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


# ------------------------------------------------------------- MSVC rand --
def test_msvc_rand_known_answers():
    """The sequence MSVC's CRT gives after srand(1) (its default seed)."""
    r = ingest.MsvcRand(1)
    assert [r() for _ in range(10)] == [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464]


# ----------------------------------------------------------- numbers --
@pytest.mark.parametrize("tok", ["0.583850", "-0.341214", "2.155820", "1", "-7", ".5", "-.25", "1e3", "2.5E-2",
                                 "123456.789", "0.000012345678", "+3.75"])
def test_number_parsing_matches_correct_rounding_on_short_decimals(tok):
    assert ingest.parse_number(tok) == np.float32(float(tok))


def test_library_numbers_equal_the_restatement(tmp_path):
    rng = np.random.default_rng(3)
    vals = [f"{v:.6f}" for v in rng.uniform(-50, 50, 300)] + [f"{v:.9e}" for v in rng.uniform(-1, 1, 100)]
    text = "".join(f"v {vals[i]} {vals[i + 1]} {vals[i + 2]}\nvn 0 0 1\n" for i in range(0, 399, 3))
    text += "".join(f"f {k + 1}//{k + 1} {k + 2}//{k + 2} {k + 3}//{k + 3}\n" for k in range(0, 130, 3))
    path = _write(tmp_path, text)
    p_lib, n_lib = hairsim.obj_triangles(path)
    p_ref, n_ref = ingest.obj_triangles(open(path).read())
    assert np.array_equal(p_lib.view(np.uint32), p_ref.view(np.uint32))


# ----------------------------------------------------- triangulation --
def _quad(pts):
    text = "".join(f"v {x} {y} {z}\n" for x, y, z in pts) + "vn 0 0 1\n"
    text += "f " + " ".join(f"{i + 1}//1" for i in range(len(pts))) + "\n"
    return text


def test_convex_quad_is_split_0_1_2_then_0_2_3(tmp_path):
    pts = [(0, 0, 0), (2, 0, 0), (2, 1, 0), (0, 1, 0)]
    p, _ = hairsim.obj_triangles(_write(tmp_path, _quad(pts)))
    P = np.array(pts, dtype=np.float32)
    assert np.array_equal(p, np.stack([P[[0, 1, 2]], P[[0, 2, 3]]]))


def test_concave_quad_avoids_the_reflex_corner(tmp_path):
    """Reflex at corner 1 (a dart): the ear search skips (0,1,2) and the
    triangles cover the dart with the diagonal 1-3."""
    pts = [(0, 0, 0), (1, 0.4, 0), (2, 0, 0), (1, 2, 0)]
    for path in (_write(tmp_path, _quad(pts)),):
        p, _ = hairsim.obj_triangles(path)
        pr, _ = ingest.obj_triangles(open(path).read())
        assert np.array_equal(p, pr)
        P = np.array(pts, dtype=np.float32)
        assert {tuple(map(tuple, t)) for t in p} == {tuple(map(tuple, P[[1, 2, 3]])), tuple(map(tuple, P[[0, 1, 3]]))}
        # the two triangles tile the dart: areas add up, both counter-clockwise like the polygon
        def tri_area(t):
            return 0.5 * ((t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[1, 1] - t[0, 1]) * (t[2, 0] - t[0, 0]))
        poly_area = 0.5 * sum(P[i, 0] * P[(i + 1) % 4, 1] - P[(i + 1) % 4, 0] * P[i, 1] for i in range(4))
        assert math.isclose(sum(tri_area(t) for t in p), poly_area, rel_tol=1e-6)
        assert all(tri_area(t) > 0 for t in p)


def test_polygons_in_other_planes_and_negative_indices(tmp_path):
    """A pentagon in the x-z plane (projection axes follow the normal) and
    relative (negative) indices: fan-like ears, area preserved."""
    ang = np.linspace(0, 2 * np.pi, 6)[:-1]
    pts = [(math.cos(a), 0.3, math.sin(a)) for a in ang]
    text = "".join(f"v {x:.6f} {y:.6f} {z:.6f}\n" for x, y, z in pts) + "vn 0 1 0\n"
    text += "f " + " ".join(f"{i - 5}//-1" for i in range(5)) + "\n"
    path = _write(tmp_path, text)
    p, n = hairsim.obj_triangles(path)
    pr, nr = ingest.obj_triangles(open(path).read())
    assert np.array_equal(p, pr) and np.array_equal(n, nr)
    assert p.shape == (3, 3, 3)
    area = sum(0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) for t in p.astype(np.float64))
    assert math.isclose(area, 2.5 * math.sin(2 * math.pi / 5), rel_tol=1e-5)


def test_missing_file_and_missing_normals(tmp_path):
    with pytest.raises(hairsim.HairSimError) as e:
        hairsim.obj_triangles(str(tmp_path / "nope.obj"))
    assert e.value.code == hairsim.HS_ERR_NOT_FOUND
    with pytest.raises(hairsim.HairSimError) as e:
        hairsim.obj_triangles(_write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "nn.obj"))
    assert e.value.code == hairsim.HS_ERR_INVALID


# ------------------------------------------------ roots and strands --
def _mesh():
    rng = np.random.default_rng(9)
    pos = rng.uniform(-1, 1, (50, 3, 3)).astype(np.float32)
    nrm = rng.normal(0, 1, (50, 3, 3)).astype(np.float32)
    return pos, nrm


def test_roots_equal_the_restatement_and_lie_on_their_triangles():
    pos, nrm = _mesh()
    roots, rn = hairsim.sample_roots(pos, nrm, 900, seed=8)
    r2, n2 = ingest.sample_roots(pos, nrm, 900, seed=8)
    assert np.array_equal(roots.view(np.uint32), r2.view(np.uint32)) and np.array_equal(rn, n2)
    # every root is a convex combination of its triangle's corners (the
    # triangle is the one whose first normal the root carries)
    for i in range(900):
        t = int(np.flatnonzero(np.all(nrm[:, 0] == rn[i], axis=1))[0])
        A = np.vstack([pos[t].T.astype(np.float64), np.ones(3)])
        w = np.linalg.lstsq(A, np.append(roots[i].astype(np.float64), 1.0), rcond=None)[0]
        assert np.all(w > -1e-4) and abs(w.sum() - 1) < 1e-5


def test_root_triangles_follow_the_rand_sequence():
    pos, nrm = _mesh()
    _, rn = hairsim.sample_roots(pos, nrm, 20, seed=8)
    r = ingest.MsvcRand(8)
    for i in range(20):
        t = r() % 50
        r(), r()
        assert np.array_equal(rn[i], nrm[t, 0])


def test_initial_strands_closed_form():
    pos, nrm = _mesh()
    roots, rn = hairsim.sample_roots(pos, nrm, 64)
    N = 10
    a = hairsim.initial_strands(roots, rn, N)
    assert np.array_equal(a.view(np.uint32), ingest.initial_strands(roots, rn, N).view(np.uint32))
    s = a.reshape(-1, 3, N, 4)
    d = rn.astype(np.float64) + np.array([0.05, 5.0, -2.0])
    want = roots[:, None, :].astype(np.float64) + np.arange(N)[None, :, None] * (2.5 / 9.0) * d[:, None, :]
    np.testing.assert_allclose(s[:, 0, :, :3], want, atol=1e-5)
    assert np.all(s[:, 0, :, 3] == 1) and np.all(s[:, 1, :, :] == np.array([0, 0, -1, 0], dtype=np.float32))
    assert np.all(s[:, 2] == 0)


@pytest.mark.skipif(not os.path.exists(MANNEQUIN), reason="the reference's mannequin model is not present")
def test_reference_mannequin_demo_groom():
    """The reference demo's groom: 383 quads -> 766 triangles, 900 roots
    (srand(8)) and 10-point strands, library == restatement bit for bit, and
    the committed fixture holds the same roots."""
    p, n = hairsim.obj_triangles(MANNEQUIN)
    pr, nr = ingest.obj_triangles(open(MANNEQUIN).read())
    assert p.shape == (766, 3, 3)
    assert np.array_equal(p.view(np.uint32), pr.view(np.uint32)) and np.array_equal(n.view(np.uint32), nr.view(np.uint32))
    roots, rn = hairsim.sample_roots(p, n, 900, seed=8)
    r2, n2 = ingest.sample_roots(pr, nr, 900, seed=8)
    assert np.array_equal(roots.view(np.uint32), r2.view(np.uint32)) and np.array_equal(rn, n2)
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "mannequin_roots.npz"))
    assert np.array_equal(fx["roots"], roots) and np.array_equal(fx["normals"], rn)


def test_last_line_without_newline_and_crlf(tmp_path):
    """The final face line may lack a newline; CRLF line ends are accepted."""
    text = "v 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nvn 0 0 1\r\nf 1//1 2//1 3//1"
    p, n = hairsim.obj_triangles(_write(tmp_path, text))
    assert p.shape == (1, 3, 3) and np.array_equal(n[0, 0], np.array([0, 0, 1], dtype=np.float32))
