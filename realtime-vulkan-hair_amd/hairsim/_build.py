"""In-tree build of libhairsim.so for gfx950 (hipcc, no cmake).

Flags that matter for results:
  -ffp-contract=off                      no fused multiply-add: the step's
                                         arithmetic is specified op by op;
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE fp32 division and sqrt;
  -fno-gpu-flush-denormals-to-zero       keep fp32 denormals (as the CPU does);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed fp32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
OUT = os.path.join(PKG, "libhairsim.so")
OBJ = os.path.join(PKG, "_build")

SOURCES = ["hs_kernels.hip", "hs_noalloc.hip", "hs_api.cpp", "hs_host.cpp", "hs_server.cpp", "hs_ingest.cpp"]
# The step kernels, once per tile shape (strands per workgroup, points per
# segment), each in its own namespace; the shapes are hs::kShapes.
STEP_SHAPES = [(128, 16), (256, 8), (256, 12)]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _units():
    """(source, object, extra flags) of every translation unit."""
    units = [(os.path.join(CSRC, s), os.path.join(OBJ, os.path.splitext(s)[0] + ".o"), []) for s in SOURCES]
    for w, l in STEP_SHAPES:
        units.append((os.path.join(CSRC, "hs_step.hip"), os.path.join(OBJ, f"hs_step_t{w}x{l}.o"),
                      [f"-DHS_TILE_STRANDS={w}", f"-DHS_TILE_POINTS={l}", f"-DHS_SHAPE_NS=t{w}x{l}"]))
    return units


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    return r.returncode, "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    units = _units()
    jobs = max(1, min(len(units), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda u: (u, _run([hipcc()] + HIPCC_FLAGS + u[2] + ["-c", "-o", u[1], u[0]])),
                                units))
    msgs = []
    for (src, obj, extra), (rc, err) in results:
        if rc != 0:
            raise RuntimeError(f"hipcc failed ({rc}) on {os.path.basename(src)} {' '.join(extra)}:\n{err}")
        if err.strip():
            msgs.append(err)
    tmp = OUT + ".tmp"
    rc, err = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [u[1] for u in units])
    if rc != 0:
        raise RuntimeError(f"link failed ({rc}):\n{err}")
    if err.strip():
        msgs.append(err)
    if verbose and msgs:
        print("\n".join(msgs), file=sys.stderr)
    os.replace(tmp, OUT)
    shutil.rmtree(OBJ, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
