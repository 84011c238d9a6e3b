"""Builds libcardiacgpu.so (HIP, gfx950) in-tree with hipcc.

Each .hip source compiles to an object in parallel, then one shared library is
linked next to this file.  Run as ``python -m cardiacfoam_amd.build``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcardiacgpu.so")
SOURCES = ["model_fp64.hip", "model_fp32.hip", "tissue_implicit.hip", "ldu_solver.hip",
           "adjoint.hip", "ecg.hip", "general_mesh.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

# Single-precision model kernels: reciprocal-based division, hardware exp2/log2 and
# fp32 denormals flushed to zero (fp64 keeps IEEE denormals; no finite-math assumption,
# so inf/NaN propagate as in IEEE arithmetic).
FAST_F32 = [
    "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fno-math-errno", "-freciprocal-math",
    "-fapprox-func", "-fno-signed-zeros", "-fgpu-flush-denormals-to-zero",
]
EXTRA = {"model_fp32.hip": FAST_F32}


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
