"""Builds the in-tree shared library ``sqdft_amd/libsqdft_amd.so`` for gfx950.

Host setup code (``csrc/system.cpp``) is compiled with ``-ffp-contract=off`` so
that its FP64 tables are bitwise equal to the reference's unfused arithmetic.
HIP sources are compiled with hipcc for gfx950 only, with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libsqdft_amd.so")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SQDFT_ARCH", "gfx950")

HOST_SRCS = ["system.cpp"]
HIP_SRCS = ["sq_gpu.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", INC, "-I", CSRC,
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I", INC]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(BUILD, s + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle() -> str:
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", d])
    return os.path.join(d, "libsq_oracle.so")


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
