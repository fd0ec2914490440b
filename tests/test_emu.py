"""The device limb arithmetic (cft-fp256_amd/csrc/fx_*.hpp) compiled for the
CPU and diffed against the oracle: every elementwise opcode the kernels run,
all four formats, all five modes, plus the golden vectors.  This checks the
kernels' logic on machines without a GPU; the gpu tests check the compiled
kernels themselves."""

import ctypes

import numpy as np
import pytest

from conftest import oracle_elem
from test_oracle import OPS, _rand_enc


def emu_elem(emu, fmt, op, rnd, a, b, c):
    esz = {0: 4, 1: 8, 2: 16, 3: 32}[fmt]
    n = len(a) // esz
    d = np.empty(n * esz, dtype=np.uint8)
    fl = ctypes.c_uint32()
    p = lambda x: None if x is None else x.ctypes.data
    emu.emu_run(fmt, op, rnd, p(a), p(b), p(c), d.ctypes.data, n, ctypes.byref(fl))
    return d, fl.value


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_emu_matches_oracle(emu, oracle, fmt):
    rng = np.random.default_rng(100 + fmt)
    for op in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20,
               21, 22, 23, 26, 27, 30, 40, 256, 257):
        for rnd in range(5):
            a, b, c = (_rand_enc(rng, fmt, 400) for _ in range(3))
            want, wf = oracle_elem(oracle, fmt, op, rnd, a, b, c)
            got, gf = emu_elem(emu, fmt, op, rnd, a, b, c)
            assert np.array_equal(got, want), (fmt, op, rnd)
            assert gf == wf, (fmt, op, rnd)


@pytest.mark.parametrize("name,fmt", [("fp128", 2), ("fp256", 3)])
def test_emu_matches_golden(emu, golden, name, fmt):
    g = golden[name]
    for opn, op in OPS.items():
        for rnd in range(5):
            b = None if opn in ("add", "sub", "sqrt") else g["b"]
            c = g["c"] if opn in ("fma", "add", "sub") else None
            d, _ = emu_elem(emu, fmt, op, rnd, g["a"], b, c)
            assert np.array_equal(d, g[f"{opn}_r{rnd}"]), (opn, rnd)


def test_emu_workload_configs(emu, oracle):
    """Config 1 and 2 operands (incl. the near-boundary FMA cases)."""
    from cftamd import workload
    for cfg, fmt in ((1, 2), (2, 3)):
        a, b, c = workload.elementwise(cfg, 20000, seed=11)
        for rnd in range(5):
            want, wf = oracle_elem(oracle, fmt, 0, rnd, a, b, c)
            got, gf = emu_elem(emu, fmt, 0, rnd, a, b, c)
            assert np.array_equal(got, want) and gf == wf, (cfg, rnd)
