"""Shared fixtures.  `gpu` marks tests that need an MI355X; everything else
runs on the CPU (the oracle, the CPU build of the device arithmetic, the
C-ABI library's exports and argument rules)."""

import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cft-fp256_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, REPO)

ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libcftoracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libcft_ref.so")
EMU_DIR = os.path.join(REPO, "tests", "emu")
EMU_SO = os.path.join(EMU_DIR, "_build", "libfxemu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


V = ctypes.c_void_p
U32P = ctypes.POINTER(ctypes.c_uint32)


def _build_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "port"], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle():
    """The oracle restatement (oracle/sf_oracle.c, seq_oracle.c)."""
    _build_oracle()
    o = ctypes.CDLL(ORACLE_SO)
    o.oc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                         ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P]
    o.oc_run_lanes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                               ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P, V]
    o.oc_compute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V, U32P]
    o.oc_div.argtypes = [ctypes.c_int, ctypes.c_int, V, V, V, U32P]
    o.oc_sqrt.argtypes = [ctypes.c_int, ctypes.c_int, V, V, U32P]
    o.oc_round_pack.argtypes = [ctypes.c_int, ctypes.c_int, V, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, V, U32P]
    o.oc_seq_run.argtypes = [V, ctypes.c_size_t, V, V, V, V, ctypes.c_size_t, V, V, V, U32P,
                             V, V, ctypes.c_uint32, ctypes.c_int, U32P, U32P]
    return o


@pytest.fixture(scope="session")
def ref():
    """The reference's own C library, where this container built it."""
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libcft_ref.so is not built here")
    r = ctypes.CDLL(REF_SO)
    r.cft_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(V)]
    r.cft_run.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                          ctypes.c_size_t, U32P, U32P]
    r.cft_div.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, V, ctypes.c_size_t, U32P, U32P]
    r.cft_sqrt.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, ctypes.c_size_t, U32P, U32P]
    return r


@pytest.fixture(scope="session")
def emu():
    """The device limb arithmetic compiled for the CPU (tests/emu)."""
    os.makedirs(os.path.join(EMU_DIR, "_build"), exist_ok=True)
    src = os.path.join(EMU_DIR, "fx_emu.cpp")
    hdrs = [os.path.join(PKG, "csrc", f) for f in os.listdir(os.path.join(PKG, "csrc"))
            if f.endswith(".hpp")] + [os.path.join(EMU_DIR, "fx_hw.hpp"), src]
    if not os.path.exists(EMU_SO) or any(os.path.getmtime(h) > os.path.getmtime(EMU_SO)
                                         for h in hdrs):
        # built under a name of this process's own, then moved into place in
        # one step: parallel test workers may rebuild it at the same time
        tmp = f"{EMU_SO}.{os.getpid()}"
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                        "-I", EMU_DIR, "-I", os.path.join(PKG, "csrc"), src, "-o", tmp],
                       check=True)
        os.replace(tmp, EMU_SO)
    e = ctypes.CDLL(EMU_SO)
    e.emu_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                          ctypes.c_size_t, U32P]
    e.emu_hardcases.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_size_t, V, V, V, V]
    return e


def oracle_elem(oracle, fmt, op, rnd, a, b, c, threads=8):
    """Oracle answer and OR of flags for an elementwise op (256 div, 257 sqrt)."""
    esz = {0: 4, 1: 8, 2: 16, 3: 32}[fmt]
    src = a if a is not None else b
    n = len(src) // esz
    d = np.empty(n * esz, dtype=np.uint8)
    fl = ctypes.c_uint32()
    p = lambda x: None if x is None else x.ctypes.data
    oracle.oc_run(fmt, op, rnd, p(a), p(b), p(c), d.ctypes.data, n, 0, threads,
                  ctypes.byref(fl))
    return d, fl.value


def oracle_flags_each(oracle, fmt, op, rnd, a, b, c):
    """Per-element flags from the oracle (slow; small n)."""
    esz = {0: 4, 1: 8, 2: 16, 3: 32}[fmt]
    n = len(a) // esz
    out = np.zeros(n, dtype=np.uint8)
    d = np.empty(esz, dtype=np.uint8)
    f = ctypes.c_uint32()
    for i in range(n):
        s = slice(i * esz, (i + 1) * esz)
        if op == 256:
            oracle.oc_div(fmt, rnd, a[s].ctypes.data, b[s].ctypes.data, d.ctypes.data,
                          ctypes.byref(f))
        elif op == 257:
            oracle.oc_sqrt(fmt, rnd, a[s].ctypes.data, d.ctypes.data, ctypes.byref(f))
        else:
            oracle.oc_compute(fmt, op, rnd, a[s].ctypes.data,
                              None if b is None else b[s].ctypes.data,
                              None if c is None else c[s].ctypes.data, d.ctypes.data,
                              ctypes.byref(f))
        out[i] = f.value
    return out


@pytest.fixture(scope="session")
def golden():
    g = {}
    for name in ("fp128", "fp256"):
        with np.load(os.path.join(REPO, "tests", "golden", f"elem_{name}.npz")) as z:
            g[name] = {k: z[k] for k in z.files}
    return g


@pytest.fixture(scope="session")
def device():
    """One MI355X handle for the GPU tests (fails loudly without the HIP lib)."""
    from cftamd import Device
    d = Device("hip:0")
    yield d
    d.close()


@pytest.fixture(scope="session")
def deep_device():
    """A handle at scratch depth 2048 (Lorenz-96 rings past 255 slots)."""
    from cftamd import Device
    d = Device("hip:0", scratch_depth=2048)
    yield d
    d.close()
