"""Negative controls, on the CPU: the checks can go red.

Each plant is a one-line mistake compiled into a copy of the oracle
(oracle/sf_oracle.c built with -DORACLE_PLANT=k into a temporary
directory): RMM rounding as RNE, UNDERFLOW dropped, RNE ties to odd,
tininess decided before rounding.  The comparison the parity tests rely on
- results and per-element flags against the golden vectors the reference's
own library wrote - must find each of them, and must find nothing in the
unplanted build.  The device arithmetic's CPU build (tests/emu) disagrees
with every planted copy too, so a GPU parity test against a wrong oracle
would fail rather than pass."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import EMU_DIR, REPO, U32P, V, oracle_elem, oracle_flags_each

OPS = {"fma": 0, "add": 1, "mul": 3, "div": 256, "sqrt": 257}
PLANTS = {1: "rmm-as-rne", 2: "drop-underflow", 3: "tie-to-odd", 4: "tininess-before-rounding"}


def build_oracle(tmp_path, plant):
    so = tmp_path / f"oracle_plant{plant}.so"
    src = [os.path.join(REPO, "oracle", f) for f in ("sf_oracle.c", "seq_oracle.c",
                                                     "rd_oracle.c")]
    subprocess.run(["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-pthread",
                    f"-DORACLE_PLANT={plant}", "-o", str(so)] + src, check=True)
    o = ctypes.CDLL(str(so))
    o.oc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                         ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P]
    o.oc_compute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V, U32P]
    o.oc_div.argtypes = [ctypes.c_int, ctypes.c_int, V, V, V, U32P]
    o.oc_sqrt.argtypes = [ctypes.c_int, ctypes.c_int, V, V, U32P]
    return o


def golden_misses(o, golden):
    """Elements whose bits or flags differ from the golden vectors, by
    operation and mode."""
    miss = {}
    for name, fmt in (("fp128", 2), ("fp256", 3)):
        g = golden[name]
        for opn, op in OPS.items():
            for rnd in range(5):
                b = None if opn in ("add", "sqrt") else g["b"]
                c = g["c"] if opn in ("fma", "add") else None
                d, _ = oracle_elem(o, fmt, op, rnd, g["a"], b, c)
                esz = 16 if fmt == 2 else 32
                bits = (d.reshape(-1, esz) != g[f"{opn}_r{rnd}"].reshape(-1, esz)).any(axis=1)
                fl = oracle_flags_each(o, fmt, op, rnd, g["a"], b, c)
                bad = int(np.count_nonzero(bits | (fl != g[f"{opn}_r{rnd}_flags"])))
                if bad:
                    miss[(name, opn, rnd)] = bad
    return miss


def closed_forms_fail(o):
    """Whether the oracle's closed-form pins (test_oracle.test_closed_forms:
    exact products, ties, overflow, the tininess boundary, cancellation,
    NaN rules) reject this build."""
    from test_oracle import test_closed_forms
    try:
        test_closed_forms(o)
    except AssertionError:
        return True
    return False


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pins_find_the_plant(tmp_path, golden, plant):
    o = build_oracle(tmp_path, plant)
    miss = golden_misses(o, golden)
    if plant == 1:                       # only round-to-nearest-away moves
        assert {k[2] for k in miss} == {4}, miss
    if plant == 4:
        # the golden families hold no product that rounds up to 2^emin from
        # below, so only the closed-form tininess boundary sees this one
        assert closed_forms_fail(o)
        return
    assert sum(miss.values()) > 0, PLANTS[plant]


def test_unplanted_build_is_clean(tmp_path, golden):
    o = build_oracle(tmp_path, 0)
    assert golden_misses(o, golden) == {}
    assert not closed_forms_fail(o)


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_device_arithmetic_disagrees_with_a_planted_oracle(tmp_path, golden, emu, plant):
    o = build_oracle(tmp_path, plant)
    g = golden["fp256"]
    # the golden operands, and the tininess boundary: the largest subnormal
    # times 1 + 2^-236 (rounds up to 2^emin under RNE: not tiny)
    B = 262143
    sub = ((1 << 236) - 1).to_bytes(32, "little")
    onep = ((B << 236) | 1).to_bytes(32, "little")
    cases = [(g["a"], g["b"], g["c"]),
             tuple(np.frombuffer(x, np.uint8).copy() for x in (sub, onep, bytes(32)))]
    diffs = 0
    for a, b, c0 in cases:
        for op in (0, 3):
            for rnd in range(5):
                c = c0 if op == 0 else None
                want, wfl = oracle_elem(o, 3, op, rnd, a, b, c)
                got = np.empty_like(want)
                fl = ctypes.c_uint32()
                emu.emu_run(3, op, rnd, a.ctypes.data, b.ctypes.data,
                            None if c is None else c.ctypes.data, got.ctypes.data,
                            a.size // 32, ctypes.byref(fl))
                diffs += int(not np.array_equal(got, want)) + int(fl.value != wfl)
    if plant == 2:
        # dropping UNDERFLOW per element can leave the call's OR unchanged
        # (other elements underflow too); the per-element golden test above
        # is the check that sees it
        return
    assert diffs > 0, PLANTS[plant]
