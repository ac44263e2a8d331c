"""Builds librbw.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librbw.so")

# -ffp-contract=off keeps a*b+c as two rounded operations (the reference's
# arithmetic); division and sqrt stay IEEE (hipcc's default).  Packed FP32
# instructions are disabled.
HIPCC_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
               "-Wno-unused-variable", "-Wno-unused-result"]


SOURCES = ("tracer.hip", "batch.hip", "scene.hip", "guiding.hip", "bvh.cpp", "image.cpp")
HOST_LIB = os.path.join(HERE, "librbw_host.so")
TEST_BIN = os.path.join(HERE, "..", "tests", "cpp", "build", "cornell_integrator")


def build_host(verbose=False):
    """The rainbow-style C++ host interface (librbw_host.so) and its C++ test driver."""
    cxx = os.environ.get("CXX", "g++")
    host = [os.path.join(HERE, "host", f) for f in ("rainbow_host.cpp", "scene_loader.cpp")]
    cmd = [cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-o", HOST_LIB + ".tmp", *host, "-L" + HERE, "-lrbw",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.makedirs(os.path.dirname(TEST_BIN), exist_ok=True)
    src = os.path.join(HERE, "..", "tests", "cpp", "cornell_integrator.cpp")
    cmd = [cxx, "-std=c++17", "-O2", "-o", TEST_BIN, src, "-L" + HERE, "-lrbw_host", "-lrbw",
           "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(TEST_BIN)).replace("..", "$ORIGIN/..", 1)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, f) for f in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-shared", "-o", tmp, *srcs, "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    build_host(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
