"""Builds libmaille_hip.so in-tree: the host pattern code with g++, the device code with hipcc
for gfx950, linked into one shared library next to this file."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmaille_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["msh.cpp", "pattern.cpp", "plan.cpp"]
DEVICE_SOURCES = ["solver.hip"]
HEADERS = ["kernels.hpp", "kernels.hip", "msh.hpp", "pattern.hpp", "plan.hpp"]

CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
            "-Wno-unused-parameter"]
# Packed FP32 instructions are disabled (the device code is written for FP64 and scalar
# FP32); -ffp-contract is left to the per-kernel pragmas.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-parameter", "-Wno-unused-command-line-argument"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
          "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
