"""Builds libb3g.so (the HIP library behind include/b3g.h) in-tree for gfx950."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libb3g.so")
SOURCES = ["b3g_api.hip", "b3g_setup.hip", "b3g_kernels.hip", "b3g_raycast.hip", "b3g_sdf.hip", "b3g_static.hip", "b3g_be.hip"]

# -ffp-contract=off and IEEE division/sqrt keep every operation rounded as the
# reference rounds it; packed FP32 instructions are disabled.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [hipcc(), *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


CONSUMER_SRC = os.path.join(ROOT, "tests", "native", "b3g_consumer.cpp")
CONSUMER = os.path.join(ROOT, "tests", "native", "_build", "b3g_consumer")


def build_consumer(verbose: bool = False) -> str:
    """The C++ test program that calls libb3g through include/b3g.h only (host compiler, linked against the
    library built above)."""
    os.makedirs(os.path.dirname(CONSUMER), exist_ok=True)
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), CONSUMER_SRC,
           "-L", HERE, "-lb3g", "-Wl,-rpath,$ORIGIN/../../../bounce_softbody_amd", "-o", CONSUMER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CONSUMER + ".tmp", CONSUMER)
    return CONSUMER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_consumer(verbose="-v" in sys.argv)
    print(LIB)
