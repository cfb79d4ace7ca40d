"""Builds libpfhip.so (HIP kernels for gfx950 + the C host driver) in-tree.

    python -m parflow_amd.build          # or parflow_amd.build.build()

hipcc compiles the .hip sources for gfx950 only; gcc compiles the C host
driver; hipcc links the shared library next to this file.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfhip.so")

HIP_SOURCES = ["core.hip", "mg.hip"]
C_SOURCES = ["kinsol_host.c", "ensemble.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             # no packed FP32 instructions: the kernels are written for the
             # scalar FP32/FP64 pipes and must not be rewritten into them
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function",
           "-I/opt/rocm/include"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
