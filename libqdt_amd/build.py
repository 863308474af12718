"""Builds libqdt_amd/_lib/libqdt_amd.so for gfx950 with hipcc (in-tree).

The library is plain C ABI (include/qdt.h, include/qdt_gpu.h) over HIP
kernels; Python loads it with ctypes.  Device code is compiled without FMA
contraction (the predicate filters' error bounds assume separate roundings)
and without packed-FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib", "libqdt_amd.so")

SOURCES = ["engine.hip", "qdt_capi.cpp", "qdt_serve.cpp", "mesh_edit.cpp", "tuner.cpp", "noalloc.hip"]
HEADERS = ["engine.hpp", "mesh_edit.hpp", "predicates.hpp", "primitives.hpp", "quadedge.hpp", "seam.hpp", "tuner.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(SRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print(OUT)
