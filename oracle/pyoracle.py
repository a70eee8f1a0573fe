"""Python loader of the CPU oracle (oracle/libmps_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker or the timed CPU
baseline.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mps_cuda_amd import ffi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmps_oracle.so")
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        ffi._declare(lib, "orc_")
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_set_threads.restype = C.c_int
        lib.orc_get_edge.argtypes = [C.c_void_p, C.c_void_p]
        lib.orc_get_edge.restype = C.c_int
        _LIB = lib
    return _LIB


def set_threads(n: int) -> int:
    """Sets the OpenMP thread count (n <= 0 keeps it); returns the count in use."""
    return load().orc_set_threads(n)


class Oracle(ffi.Handle):
    def __init__(self, params, scene):
        super().__init__(params, scene, lib=load(), prefix="orc_")

    def edge(self) -> np.ndarray:
        e = np.zeros(self.n, np.int8)
        self._check(self.lib.orc_get_edge(self._h, e.ctypes.data))
        return e
