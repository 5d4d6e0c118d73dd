"""Build libgfn1_amd.so (HIP kernels + C++ host + C ABI) for gfx950 in-tree.

Usage: python gfn1-rs_amd/build.py   (or __graft_entry__.build())
Output: gfn1-rs_amd/lib/libgfn1_amd.so
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIB_DIR, "libgfn1_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("GFN1_OFFLOAD_ARCH", "gfx950")

SOURCES = ["model.cpp", "capi.cpp", "engine.hip", "kernels.hip", "eigen.hip", "guess.hip", "cholesky.hip", "density.hip", "tridiag_large.hip", "periodic.hip", "optimizer.cpp", "hessian.cpp", "md.cpp"]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROCM, "include"),
    # packed FP32 arithmetic off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=" + ARCH, "-x", "hip"] + COMMON + \
          ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB] + objs + \
          ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
