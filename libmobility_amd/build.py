"""Build the gfx950 shared library libmobility_amd/lib/liblmamd.so with hipcc.

Usage: python -m libmobility_amd.build   (or __graft_entry__.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "liblmamd.so")
OBJ_DIR = os.path.join(HERE, "_build")

SOURCES = ["nbody_kernels.hip", "lanczos_kernels.hip", "precond_kernels.hip",
           "coarse_kernels.hip", "noalloc.hip", "coarse.cpp", "precond.cpp", "solver.cpp", "self_solver.cpp", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LM_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are switched off for every device compile (the
# compiler may not form v_pk_{add,mul,fma}_f32 on its own).
COMMON = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-result",
    f"-I{os.path.join(HERE, '..', 'include')}",
]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[libmobility_amd] built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
