"""The hard-case generator of the differential stream (cft-fp256_amd/csrc/
hardcase.hpp), host build (tests/emu), checked on the CPU against the oracle:
that a case depends only on (seed, index, format, op), and that each family
makes the cases it is named for (exact cancellations exact, ties tied,
overflow and underflow edges raising those flags).  The GPU build is held to
this one bit for bit in tests/test_diffstream_gpu.py.  Also the oracle's
per-element flags (oc_run_lanes), which the stream's flag check rests on."""

import ctypes

import numpy as np
import pytest

from conftest import oracle_flags_each

OPS = {"fma": 0, "add": 1, "sub": 2, "mul": 3, "div": 256, "sqrt": 257}
INVALID, OVERFLOW, UNDERFLOW, INEXACT = 1, 4, 8, 16


def gen(emu, fmt, op, seed, base, n):
    es = 4 << fmt
    a, b, c = (np.zeros((n, es), np.uint8) for _ in range(3))
    fam = np.zeros(n, np.uint8)
    assert emu.emu_hardcases(fmt, op, seed, base, n, a.ctypes.data, b.ctypes.data,
                             c.ctypes.data, fam.ctypes.data) == 0
    return a, b, c, fam


def lanes(oracle, fmt, op, rnd, a, b, c):
    n, es = a.shape
    d = np.empty((n, es), np.uint8)
    lf = np.empty(n, np.uint8)
    fl = ctypes.c_uint32()
    rb = b.ctypes.data if op in (0, 3, 256) else None
    rc = c.ctypes.data if op in (0, 1, 2) else None
    oracle.oc_run_lanes(fmt, op, rnd, a.ctypes.data, rb, rc, d.ctypes.data, n, 0, 8,
                        ctypes.byref(fl), lf.ctypes.data)
    return d, lf, fl.value


@pytest.mark.parametrize("fmt", [0, 3])
@pytest.mark.parametrize("op", [0, 1, 256, 257])
def test_counter_based(emu, fmt, op):
    """Case g is the same from any starting index, differs by seed and op."""
    a, b, c, fam = gen(emu, fmt, op, 7, 1000, 300)
    a2, b2, c2, fam2 = gen(emu, fmt, op, 7, 1100, 200)
    assert (a[100:] == a2).all() and (b[100:] == b2).all() and (c[100:] == c2).all()
    assert (fam == (np.arange(1000, 1300) % 14)).all()
    a3, _, _, _ = gen(emu, fmt, op, 8, 1000, 300)
    assert (a3 != a).any(axis=1).mean() > 0.9
    if op != 0:
        a4, _, _, _ = gen(emu, fmt, 0, 7, 1000, 300)
        assert (a4 != a).any(axis=1).mean() > 0.3


def test_refuses_unknown(emu):
    z = np.zeros(64, np.uint8)
    assert emu.emu_hardcases(4, 0, 1, 0, 1, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                             None) == -1
    assert emu.emu_hardcases(3, 5, 1, 0, 1, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                             None) == -1


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_families_make_their_cases(emu, oracle, fmt):
    n = 14 * 600
    a, b, c, fam = gen(emu, fmt, 0, 11, 0, n)
    d_ne, lf_ne, _ = lanes(oracle, fmt, 0, 0, a, b, c)
    d_mm, _, _ = lanes(oracle, fmt, 0, 4, a, b, c)
    tie = (d_ne != d_mm).any(axis=1)      # RNE and RMM differ only on a midpoint
    frac = lambda f, m: m[fam == f].mean()
    # 5: c = -(a*b) of an exact product, or -(a*b rounded): the FMA is exact
    assert (lf_ne[fam == 5] == 0).all()
    # 6: a*b exact, c a half ulp (or just off it): inexact, and many exact ties
    assert (lf_ne[fam == 6] & INEXACT).all()
    assert frac(6, tie) > 0.05
    others = np.isin(fam, [1, 2, 3, 7, 11, 13])
    assert tie[others].mean() < 0.02
    # the edges raise their flags
    assert frac(2, (lf_ne & UNDERFLOW) != 0) > 0.3
    assert frac(9, (lf_ne & UNDERFLOW) != 0) > 0.2
    assert frac(3, (lf_ne & OVERFLOW) != 0) > 0.2
    assert frac(12, (lf_ne & OVERFLOW) != 0) > 0.1
    assert frac(8, (lf_ne & INVALID) != 0) > 0.1
    # 4, 9, 10: c within 8 ulps of -(a*b rounded): the sum is small against a*b
    # (the exact ones dominate)
    assert frac(4, lf_ne == 0) > 0.5
    # the sums of the ADD stream are the same classes: family 5's sums are 0
    a1, _, c1, fam1 = gen(emu, fmt, 1, 11, 0, n)
    d1, lf1, _ = lanes(oracle, fmt, 1, 0, a1, a1, c1)
    z = d1[fam1 == 5].copy()
    z[:, -1] &= 0x7F
    assert (z == 0).all() and (lf1[fam1 == 5] == 0).all()


@pytest.mark.parametrize("fmt", [0, 3])
@pytest.mark.parametrize("op", [0, 1, 3, 256, 257])
def test_oracle_lane_flags(emu, oracle, fmt, op):
    """oc_run_lanes: each element's flags are oc_compute's (oc_div's,
    oc_sqrt's) for it alone, their OR is the call's, the bits oc_run's."""
    a, b, c, _ = gen(emu, fmt, op, 3, 0, 14 * 20)
    for rnd in (0, 2, 4):
        d, lf, fl = lanes(oracle, fmt, op, rnd, a, b, c)
        each = oracle_flags_each(oracle, fmt, op, rnd, a.reshape(-1),
                                 b.reshape(-1) if op in (0, 3, 256) else None,
                                 c.reshape(-1) if op in (0, 1, 2) else None)
        assert (lf == each).all()
        assert fl == int(np.bitwise_or.reduce(lf))
        d2 = np.empty_like(d)
        f2 = ctypes.c_uint32()
        oracle.oc_run(fmt, op, rnd, a.ctypes.data, b.ctypes.data if op in (0, 3, 256) else None,
                      c.ctypes.data if op in (0, 1, 2) else None, d2.ctypes.data, a.shape[0], 0,
                      1, ctypes.byref(f2))
        assert (d == d2).all() and f2.value == fl


@pytest.mark.parametrize("fmt", [2, 3])
def test_reference_agrees_on_hard_cases(emu, oracle, ref, fmt):
    """The oracle and the reference's CPU library on the stream's cases: the
    same bits (a CPU-only preview of what the GPU stream compares)."""
    V = ctypes.c_void_p
    dev = V()
    assert ref.cft_open(None, 0, ctypes.byref(dev)) == 0
    try:
        for op in (0, 1, 2, 3, 256, 257):
            a, b, c, _ = gen(emu, fmt, op, 5, 0, 14 * 40)
            for rnd in range(5):
                d, _, fl = lanes(oracle, fmt, op, rnd, a, b, c)
                r = np.empty_like(d)
                f = ctypes.c_uint32()
                if op == 256:
                    st = ref.cft_div(dev, fmt, rnd, a.ctypes.data, b.ctypes.data, r.ctypes.data,
                                     a.shape[0], ctypes.byref(f), None)
                elif op == 257:
                    st = ref.cft_sqrt(dev, fmt, rnd, a.ctypes.data, r.ctypes.data, a.shape[0],
                                      ctypes.byref(f), None)
                else:
                    st = ref.cft_run(dev, op, fmt, rnd, a.ctypes.data,
                                     b.ctypes.data if op in (0, 3) else None,
                                     c.ctypes.data if op in (0, 1, 2) else None, r.ctypes.data,
                                     a.shape[0], ctypes.byref(f), None)
                assert st == 0
                assert (r == d).all(), (op, rnd)
                assert f.value == fl
    finally:
        ref.cft_close(dev)


CLANG = "/opt/rocm/llvm/bin/clang++"


def test_generator_same_under_two_compilers(emu, tmp_path):
    """The order of evaluation of a call's arguments is unspecified, and g++
    and clang (the GPU build's compiler) choose differently: the generator
    must make the same cases under both, so a second host build with clang
    is held to the first."""
    import os
    import subprocess
    from conftest import EMU_DIR, PKG
    if not os.path.exists(CLANG):
        pytest.skip("no clang++ here")
    so = str(tmp_path / "libfxemu_clang.so")
    subprocess.run([CLANG, "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                    "-Wno-pass-failed",
                    "-I", EMU_DIR, "-I", os.path.join(PKG, "csrc"),
                    os.path.join(EMU_DIR, "fx_emu.cpp"), "-o", so], check=True)
    other = ctypes.CDLL(so)
    other.emu_hardcases.argtypes = emu.emu_hardcases.argtypes
    for fmt in range(4):
        for op in OPS.values():
            x = gen(emu, fmt, op, 31, 77, 14 * 50)
            y = gen(other, fmt, op, 31, 77, 14 * 50)
            for u, v in zip(x, y):
                assert (u == v).all(), (fmt, op)
