"""Builds the gfx950 shared library ``mps_cuda_amd/libmps_hip.so`` in-tree.

Run as ``python -m mps_cuda_amd.build`` or through ``__graft_entry__.build()``.
The HIP code is compiled with -ffp-contract=off (no fused multiply-add the
CPU restatement would not do) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmps_hip.so")

SOURCES = ["mps_api.hip", "mps_host.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


DRIVER = os.path.join(HERE, "mps_run")


def build_driver(verbose: bool = False) -> str:
    """The command-line driver (host C++ over the C ABI, linked to the library)."""
    tmp = DRIVER + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "mps_run.cpp"), "-L", HERE, "-lmps_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building mps_run failed")
    os.replace(tmp, DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmps_hip.so")
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
