"""Builds wasl_amd/libwasl_hip.so in-tree for gfx950.

Host setup (setup.cpp) is compiled with g++ without FMA contraction so the
reference-element tables hold the same doubles as the reference's own
evaluations; device code (engine.hip) is compiled with hipcc for gfx950.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwasl_hip.so")
ARCH = os.environ.get("WASL_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 arithmetic stays off on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    setup_o = os.path.join(OUT, "setup.o")
    engine_o = os.path.join(OUT, "engine.o")
    _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "setup.cpp"), "-o", setup_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "engine.hip"), "-o", engine_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", engine_o, setup_o, "-o", tmp,
          "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
