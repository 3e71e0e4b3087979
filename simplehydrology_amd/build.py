"""Build recipe for the gfx950 shared library ``simplehydrology_amd/libshy.so``.

hipcc compiles the HIP sources for gfx950 only.  ``-ffp-contract=off`` keeps
every float rounding where the reference's source puts it (no fused
multiply-add), which the bit-exact parity tests rely on; packed FP32
instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libshy.so")
SOURCES = ["shy_engine.hip", "shy_nn.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-command-line-argument",
]


def build_library(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, LIB)
    return LIB


DEMO = os.path.join(HERE, "world_demo")


def build_demo() -> str:
    """The C++ facade example (include/shy_world.hpp) linked against libshy.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", DEMO + ".tmp",
           os.path.join(CSRC, "world_demo.cpp"), "-L" + HERE, "-lshy", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(DEMO + ".tmp", DEMO)
    return DEMO


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_demo())
