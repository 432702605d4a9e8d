"""Compile the gfx950 HIP engine into ``ziran2019_amd/lib/libzr_mpm.so``.

Run ``python -m ziran2019_amd.build`` (or ``__graft_entry__.build()``).  The
library is built in-tree so it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libzr_mpm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # LDS / global FP atomics as native instructions rather than CAS loops
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
]


HOST = os.path.join(HERE, "host")
HOST_EXE = os.path.join(LIBDIR, "zr_host_step")


def build_host(verbose=False):
    """The C++ host side (MpmSimulationGpu) and its driver program, linked
    against libzr_mpm.so with g++."""
    tmp = HOST_EXE + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(HERE, "..", "include"), "-I", HOST,
           os.path.join(HOST, "MpmSimulationGpu.cpp"), os.path.join(HOST, "zr_host_step.cpp"),
           "-L", LIBDIR, "-lzr_mpm", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_EXE)
    return HOST_EXE


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(CSRC, "zr_engine.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
