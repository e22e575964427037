"""Writes tests/golden/mannequin_roots.npz: the reference demo's 900 hair roots
and root normals (src/Strand.cpp:26-107: srand(8) over the triangulated
src/models/mannequin_segment.obj), computed by the Python restatement in
oracle/ingest_oracle.py.  Run here, where the reference tree is present; the GPU
tests use the fixture."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import ingest_oracle as ingest  # noqa: E402

OBJ = "/root/reference/src/models/mannequin_segment.obj"

if __name__ == "__main__":
    pos, nrm = ingest.obj_triangles(open(OBJ).read())
    roots, rn = ingest.sample_roots(pos, nrm, 900, seed=8)
    np.savez_compressed(os.path.join(HERE, "mannequin_roots.npz"), roots=roots, normals=rn)
    print(pos.shape[0], "triangles;", roots.shape[0], "roots")
