"""Builds libbemfmm.so in-tree for gfx950.

Host code (plan.cpp, operators.cpp, schwarz.cpp, precond.cpp: the plan, its float128
operators, the Schwarz blocks and the ILU(0) matrix; bemfmm.cpp: contexts, launch sequences, GMRES and
the C ABI over the HIP runtime API; server.cpp: the request loop over the C ABI)
is compiled with g++ (OpenMP via libgomp, __float128 via libquadmath, threads);
the device kernels and their launchers (kernels.hip, tree.hip) with hipcc for gfx950; the
objects are linked with hipcc.
Usage: python -m bem_for_nonlinear_waves_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbemfmm.so")
ARCH = os.environ.get("BEMFMM_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# host files that call the HIP runtime API (bemfmm.cpp) are plain C++ against its header
HIP_API_FLAGS = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-pthread", "-march=x86-64-v2", "-Wall", "-Wno-unused-function"]


def _gcc_libdir() -> str:
    out = subprocess.run(["g++", "-print-file-name=libgomp.so"], capture_output=True, text=True, check=True).stdout.strip()
    return os.path.dirname(out)


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in ("plan.cpp", "operators.cpp", "schwarz.cpp", "precond.cpp", "server.cpp", "bemfmm.cpp"):
        obj = os.path.join(OUT_DIR, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, *HIP_API_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(OUT_DIR, "kernels.o")
    _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "kernels.hip"), "-o", obj])
    objs.append(obj)
    # the tree kernels round their box arithmetic as the host plan does: no contraction into FMAs
    obj = os.path.join(OUT_DIR, "tree.o")
    _run([hipcc, *HIP_FLAGS, "-ffp-contract=off", "-c", os.path.join(CSRC, "tree.hip"), "-o", obj])
    objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{_gcc_libdir()}",
          f"-Wl,-rpath,{_gcc_libdir()}", "-lgomp", "-lquadmath", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
