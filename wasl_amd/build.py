"""Builds wasl_amd/libwasl_hip.so and wasl_amd/libwasl_hip_case.so in-tree.

Host setup (setup.cpp) is compiled with g++ without FMA contraction so the
reference-element tables hold the same doubles as the reference's own
evaluations; device code (engine.hip) is compiled with hipcc for gfx950.
The case layer (case_io.cpp, case_runner.cpp: the coarse calls, the runner
and the HDF5 formats) is plain C++ over the engine's C ABI, linked against
the HDF5 C library.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwasl_hip.so")
CASE_LIB = os.path.join(HERE, "libwasl_hip_case.so")
HDF5 = os.environ.get("WASL_HIP_HDF5", "/opt/conda")
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
    build_case()
    return LIB


def build_case() -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for name in ("case_io", "case_runner"):
        o = os.path.join(OUT, name + ".o")
        _run(["g++", "-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", f"-I{HDF5}/include",
              "-c", os.path.join(CSRC, name + ".cpp"), "-o", o])
        objs.append(o)
    # A directory holding only the HDF5 library: putting HDF5's own lib
    # directory on the run path would also put its older libstdc++ there.
    h5dir = os.path.join(HERE, "_hdf5")
    os.makedirs(h5dir, exist_ok=True)
    link = os.path.join(h5dir, "libhdf5.so.103")
    if not os.path.lexists(link):
        os.symlink(os.path.join(HDF5, "lib", "libhdf5.so.103"), link)
    tmp = CASE_LIB + ".tmp"
    _run(["g++", "-shared", "-fPIC", *objs, "-o", tmp, f"-L{HERE}", "-lwasl_hip",
          f"-L{h5dir}", "-l:libhdf5.so.103", "-Wl,-rpath,$ORIGIN:$ORIGIN/_hdf5"])
    os.replace(tmp, CASE_LIB)
    return CASE_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
