"""Seeded synthetic operands for the benchmark configurations (libcftwork.so,
built from csrc/workload.c).  Element i is a pure function of (seed, i)."""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libcftwork.so")
_W = None


def _w():
    global _W
    if _W is None:
        if not os.path.exists(_PATH):
            raise OSError(f"{_PATH} is not built; run make -C cft-fp256_amd")
        _W = ctypes.CDLL(_PATH)
        for name in ("cftw_elem_c1", "cftw_elem_c2"):
            fn = getattr(_W, name)
            fn.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t,
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_int]
            fn.restype = ctypes.c_int
    return _W


def elementwise(config, n, seed=0, first=0, threads=8, out=None):
    """(a, b, c) uint8 arrays for config 1 (binary128) or 2 (binary256)."""
    esz = 16 if config == 1 else 32
    if out is None:
        out = tuple(np.empty(n * esz, dtype=np.uint8) for _ in range(3))
    a, b, c = out
    fn = _w().cftw_elem_c1 if config == 1 else _w().cftw_elem_c2
    fn(seed, first, n, a.ctypes.data, b.ctypes.data, c.ctypes.data, threads)
    return a, b, c
