"""CPU property test of the mesh collider's host preprocessing
(realtime-vulkan-hair_amd/csrc/hs_mesh_build.h; the mesh collider has no
reference counterpart).  tests/mesh_build_check.cpp is built with g++ against
the header and checks, on the head meshes of the workloads: candidate records
hold bit copies of their triangle's vertices; the bounding radius covers every
vertex around the centre the kernel forms in binary32; keys ascend per cell and
bound the cell-centre distance from below; every triangle within the margin of
a sampled point near the surface is in that point's cell list."""
import inspect
import os
import shutil
import subprocess

import numpy as np
import pytest

from hairsim import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "realtime-vulkan-hair_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path_factory.mktemp("mesh") / "mesh_build_check")
    r = subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "mesh_build_check.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run_check(checker, tmp_path, verts, tris, dims, margin, samples):
    path = tmp_path / "mesh.bin"
    with open(path, "wb") as f:
        f.write(np.array([len(verts), len(tris), *dims], dtype=np.uint32).tobytes())
        f.write(np.array([margin], dtype=np.float32).tobytes())
        f.write(np.ascontiguousarray(verts, dtype=np.float32).tobytes())
        f.write(np.ascontiguousarray(tris, dtype=np.uint32).tobytes())
    r = subprocess.run([checker, str(path), str(samples)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, entries, checked = r.stdout.split()
    assert ok == "ok" and int(entries) > 0 and int(checked) > samples // 4
    return int(entries)


def test_small_icosphere(checker, tmp_path):
    verts, tris = scenes.icosphere(3, (1.0, 1.2, 1.0))
    run_check(checker, tmp_path, verts, tris, (24, 28, 24), 0.05, 2000)


def test_head_mesh_of_the_large_workload(checker, tmp_path):
    # the head of the 1M x 64 workload: the scene's own defaults
    p = inspect.signature(scenes.mesh_head_scene).parameters
    verts, tris = scenes.icosphere(5, p["scale"].default)
    run_check(checker, tmp_path, verts, tris, p["mesh_grid"].default, p["margin"].default, 400)
