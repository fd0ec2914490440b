"""The oracle's reductions (oracle/rd_oracle.c) pinned on the CPU: against
the reference's own compiled library (cft_reduce / cft_reduce_seg on its
software backend) over every opcode, format and rounding attribute and the
sizes where the tree's shape shows, and against what the contract fixes by
itself - the shape (the left child the largest power of two strictly inside
the range) through sums whose rounding depends on it, the identities of an
empty array, a lone element returned verbatim, and 754-2019 9.4's infinity
ahead of NaN for sumSquare and sumAbs."""

import ctypes
import struct

import numpy as np
import pytest

from cftamd import programs as P

V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)
ESZ = {0: 4, 1: 8, 2: 16, 3: 32}
OPS = {"sum": 24, "dot": 25, "sumsq": 28, "sumabs": 29, "maxall": 31}


@pytest.fixture(scope="module")
def oc(oracle):
    oracle.oc_reduce.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                                 ctypes.c_size_t, U32P]
    oracle.oc_reduce_seg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                                     ctypes.c_size_t, ctypes.c_size_t, U32P]
    return oracle


@pytest.fixture(scope="module")
def refdev(ref):
    ref.cft_reduce.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                               ctypes.c_size_t, U32P, U32P]
    ref.cft_reduce_seg.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                                   ctypes.c_size_t, ctypes.c_size_t, U32P, U32P]
    h = V()
    assert ref.cft_open(None, 0, ctypes.byref(h)) == 0
    return ref, h


def red_values(rng, fmt, n, special=0.03):
    """Values a sum keeps meaningful: mostly normal numbers within a few
    binades of 1 (either sign), some far apart, and a few specials."""
    width, p, ew = P.FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    esz = width // 8
    top = lambda s, e, f: (s << (width - 1)) | (e << (p - 1)) | f
    out = np.empty(n * esz, np.uint8)
    for i in range(n):
        r = rng.random()
        s = int(rng.integers(0, 2))
        if r < special:
            v = [0, 1 << (width - 1), top(0, (1 << ew) - 1, 0), top(1, (1 << ew) - 1, 0),
                 top(0, (1 << ew) - 1, 1 << (p - 2)), top(0, (1 << ew) - 1, 1),
                 top(s, 0, 1)][int(rng.integers(0, 7))]
        else:
            spread = 6 if r < 0.8 else min(bias - 2, 3 * p)
            e = bias + int(rng.integers(-spread, spread + 1))
            f = int.from_bytes(rng.bytes(esz), "little") & ((1 << (p - 1)) - 1)
            v = top(s, e, f)
        out[i * esz:(i + 1) * esz] = np.frombuffer(v.to_bytes(esz, "little"), np.uint8)
    return out


def run_oc(oc, fmt, op, rnd, a, b, n, seg=None):
    d = np.zeros(max(1, (n // seg) if seg else 1) * ESZ[fmt], np.uint8)
    fl = ctypes.c_uint32(0xDEAD)
    pa = a.ctypes.data if a is not None else None
    pb = b.ctypes.data if b is not None else None
    if seg is None:
        assert oc.oc_reduce(fmt, op, rnd, pa, pb, d.ctypes.data, n, ctypes.byref(fl)) == 0
    else:
        assert oc.oc_reduce_seg(fmt, op, rnd, pa, pb, d.ctypes.data, n, seg,
                                ctypes.byref(fl)) == 0
    return d, fl.value


def run_ref(refdev, fmt, op, rnd, a, b, n, seg=None):
    ref, h = refdev
    d = np.zeros(max(1, (n // seg) if seg else 1) * ESZ[fmt], np.uint8)
    fl, bus = ctypes.c_uint32(), ctypes.c_uint32()
    pa = a.ctypes.data if a is not None else None
    pb = b.ctypes.data if b is not None else None
    if seg is None:
        st = ref.cft_reduce(h, op, fmt, rnd, pa, pb, d.ctypes.data, n, ctypes.byref(fl),
                            ctypes.byref(bus))
    else:
        st = ref.cft_reduce_seg(h, op, fmt, rnd, pa, pb, d.ctypes.data, n, seg,
                                ctypes.byref(fl), ctypes.byref(bus))
    assert st == 0
    return d, fl.value


SIZES = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 15, 16, 17, 31, 33, 64, 100, 129]


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_oracle_matches_the_reference(oc, refdev, fmt):
    rng = np.random.default_rng(10 + fmt)
    for op in OPS.values():
        for n in SIZES:
            rnd = int(rng.integers(0, 5))
            a = red_values(rng, fmt, max(n, 1))
            b = red_values(rng, fmt, max(n, 1))
            want = run_ref(refdev, fmt, op, rnd, a, b, n)
            got = run_oc(oc, fmt, op, rnd, a, b, n)
            assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], (op, n, rnd)


@pytest.mark.parametrize("fmt", [0, 3])
def test_oracle_matches_the_reference_segmented(oc, refdev, fmt):
    rng = np.random.default_rng(20 + fmt)
    n = 60
    for op in OPS.values():
        for seg in (1, 2, 3, 4, 5, 6, 10, 12, 15, 20, 30, 60):
            for rnd in range(5):
                a = red_values(rng, fmt, n, special=0.05)
                b = red_values(rng, fmt, n)
                want = run_ref(refdev, fmt, op, rnd, a, b, n, seg)
                got = run_oc(oc, fmt, op, rnd, a, b, n, seg)
                assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], \
                    (op, seg, rnd)


def test_oracle_matches_the_reference_on_special_mixes(oc, refdev):
    # infinities, NaNs of both kinds and zeros of both signs placed so every
    # override and every n <= 2 edge is met
    fmt = 1
    specials = [struct.pack("<d", v) for v in (float("inf"), float("-inf"), 0.0, -0.0, 1.5,
                                               -2.25)]
    specials += [struct.pack("<Q", 0x7FF8000000000000), struct.pack("<Q", 0x7FF0000000000001),
                 struct.pack("<Q", 0xFFF0000000000002)]
    rng = np.random.default_rng(5)
    for trial in range(400):
        n = int(rng.integers(0, 6))
        a = np.frombuffer(b"".join(specials[int(rng.integers(0, len(specials)))]
                                   for _ in range(max(n, 1))), np.uint8).copy()
        b = np.frombuffer(b"".join(specials[int(rng.integers(0, len(specials)))]
                                   for _ in range(max(n, 1))), np.uint8).copy()
        for op in OPS.values():
            rnd = int(rng.integers(0, 5))
            got, want = run_oc(oc, fmt, op, rnd, a, b, n), run_ref(refdev, fmt, op, rnd, a, b, n)
            assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], (trial, op, n)


def f32(x):
    return np.frombuffer(struct.pack("<" + "f" * len(x), *x), np.uint8).copy()


def as_f32(d):
    return struct.unpack("<f", d.tobytes()[:4])[0]


def test_the_shape_is_the_binary_counter_tree(oc):
    e = 2.0 ** -24
    # T(0,3) = (a0 + a1) + a2: 1 + 2^-24 ties to 1 twice
    assert as_f32(run_oc(oc, 0, 24, 0, f32([1.0, e, e]), None, 3)[0]) == 1.0
    # the other order of the same values: the small ones meet first
    assert as_f32(run_oc(oc, 0, 24, 0, f32([e, e, 1.0]), None, 3)[0]) == 1.0 + 2 * e
    # T(0,5) = T(0,4) + a4 = ((1 + 0) + (e + 0)) + e: 1 + e ties to 1, twice;
    # the floor-midpoint tree T(0,2) + T(2,5) = 1 + 2e would be exact
    assert as_f32(run_oc(oc, 0, 24, 0, f32([1.0, 0.0, e, 0.0, e]), None, 5)[0]) == 1.0
    # and upward rounding sees the same tree: (1 + e) -> 1 + 2e, then + e -> 1 + 4e
    assert as_f32(run_oc(oc, 0, 24, 3, f32([1.0, 0.0, e, 0.0, e]), None, 5)[0]) == 1.0 + 4 * e


def test_identities_and_lone_elements(oc):
    z = np.zeros(32, np.uint8)
    d, fl = run_oc(oc, 3, 24, 0, z, None, 0)
    assert d.tobytes() == bytes(32) and fl == 0
    d, fl = run_oc(oc, 3, 31, 0, z, None, 0)              # maxall of nothing: -inf
    assert d.tobytes() == bytes(29) + bytes([0xF0, 0xFF, 0xFF]) and fl == 0
    snan = np.frombuffer(struct.pack("<Q", 0xFFF0000000000001), np.uint8).copy()
    for op in (24, 31):                                   # verbatim, no flags
        d, fl = run_oc(oc, 1, op, 0, snan, None, 1)
        assert d.tobytes() == snan.tobytes() and fl == 0
    d, fl = run_oc(oc, 1, 29, 0, snan, None, 1)           # abs: sign cleared, no flags
    assert struct.unpack("<Q", d.tobytes())[0] == 0x7FF0000000000001 and fl == 0
    d, fl = run_oc(oc, 1, 28, 0, snan, None, 1)           # the leaf is a multiply
    assert struct.unpack("<Q", d.tobytes())[0] == 0x7FF8000000000000 and fl == 1


def test_sumsq_and_sumabs_put_infinity_ahead_of_nan(oc):
    inf, qnan = struct.pack("<d", float("inf")), struct.pack("<Q", 0x7FF8000000000000)
    snan = struct.pack("<Q", 0x7FF0000000000001)
    a = np.frombuffer(inf + qnan, np.uint8).copy()
    for op in (28, 29):
        d, fl = run_oc(oc, 1, op, 0, a, None, 2)
        assert struct.unpack("<d", d.tobytes())[0] == float("inf") and fl == 0
    d, fl = run_oc(oc, 1, 24, 0, a, None, 2)              # plain sum: NaN first
    assert struct.unpack("<Q", d.tobytes())[0] == 0x7FF8000000000000
    a = np.frombuffer(snan + inf, np.uint8).copy()
    for op in (28, 29):
        d, fl = run_oc(oc, 1, op, 0, a, None, 2)
        assert struct.unpack("<d", d.tobytes())[0] == float("inf") and fl == 1
