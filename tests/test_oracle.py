"""The oracle is pinned before anything is checked against it:

* against the reference's own compiled C library (oracle/_ref), where this
  container built it, over structured random cases;
* against the committed golden vectors, which that library wrote
  (tests/golden/make_golden.py), so the pin holds on machines without it;
* against closed forms that a plausible mistake in rounding would break.
"""

import ctypes

import numpy as np
import pytest

from conftest import oracle_elem, oracle_flags_each

OPS = {"fma": 0, "add": 1, "sub": 2, "mul": 3, "div": 256, "sqrt": 257}


@pytest.mark.parametrize("name,fmt", [("fp128", 2), ("fp256", 3)])
@pytest.mark.parametrize("opn", list(OPS))
def test_oracle_matches_golden(oracle, golden, name, fmt, opn):
    g = golden[name]
    op = OPS[opn]
    for rnd in range(5):
        b = None if opn in ("add", "sub", "sqrt") else g["b"]
        c = g["c"] if opn in ("fma", "add", "sub") else None
        d, _ = oracle_elem(oracle, fmt, op, rnd, g["a"], b, c)
        assert np.array_equal(d, g[f"{opn}_r{rnd}"]), (opn, rnd)
        fl = oracle_flags_each(oracle, fmt, op, rnd, g["a"], b, c)
        assert np.array_equal(fl, g[f"{opn}_r{rnd}_flags"]), (opn, rnd)


def _rand_enc(rng, fmt, n):
    ew, mw, w = {0: (8, 23, 32), 1: (11, 52, 64), 2: (15, 112, 128), 3: (19, 236, 256)}[fmt]
    esz, bias = w // 8, (1 << (ew - 1)) - 1
    out = []
    for _ in range(n):
        r = rng.random()
        s = int(rng.integers(0, 2))
        m = int.from_bytes(rng.bytes(esz), "little") % (1 << mw)
        if r < 0.1:
            e = 0
            m >>= int(rng.integers(0, mw))
        elif r < 0.15:
            e = (1 << ew) - 1
            m = [0, 1, 1 << (mw - 1)][int(rng.integers(0, 3))]
        elif r < 0.55:
            e = bias + int(rng.integers(-(mw + 4), mw + 4))
        else:
            e = int(rng.integers(1, (1 << ew) - 1))
        e = min(max(e, 0), (1 << ew) - 1)
        out.append(((s << (w - 1)) | (e << mw) | m).to_bytes(esz, "little"))
    return np.frombuffer(b"".join(out), dtype=np.uint8).copy()


def _next_to(rng, fmt, n):
    """(a, b) with b within four ulps of a (either side, a fifth of the b
    significands then set to all ones): a/b and b/a lie next to 1, so the
    long division's quotient digits are B-1 or 0 and its remainder's top
    limb meets the divisor's; square roots of either are next to a power of
    two's root."""
    ew, mw, w = {0: (8, 23, 32), 1: (11, 52, 64), 2: (15, 112, 128), 3: (19, 236, 256)}[fmt]
    esz, bias = w // 8, (1 << (ew - 1)) - 1
    span = min(256, 1 << (ew - 2))
    aa, bb = [], []
    for i in range(n):
        m = int.from_bytes(rng.bytes(esz), "little") % (1 << mw)
        x = ((bias + int(rng.integers(-span, span))) << mw) | m
        k = int(rng.integers(1, 5)) * (1 if i % 2 else -1)
        y = x + k
        if i % 5 == 0:
            y |= (1 << mw) - 1
        aa.append(x.to_bytes(esz, "little"))
        bb.append(y.to_bytes(esz, "little"))
    return (np.frombuffer(b"".join(aa), dtype=np.uint8).copy(),
            np.frombuffer(b"".join(bb), dtype=np.uint8).copy())


def _edges(rng, fmt, n, kind):
    """(a, b) at the thresholds where division and multiplication settle a
    result by its exponent alone.  kind "div": a subnormal dividend over a
    normal divisor whose lsb exponent is -3..3 (from 1 up, the quotient is
    P+1 or more places below 2^emin), and a normal dividend of leading
    exponent -1..4 over a subnormal divisor (from 2 up, the quotient is past
    the overflow threshold), a tenth of the pairs both subnormal.  kind
    "mul": a subnormal factor times a normal one whose lsb exponent is
    -2P-4 .. -2P+2 (at -2P-1 and below the product's top bit lies P+1 or
    more places below 2^emin).  Random signs and significands."""
    ew, mw, w = {0: (8, 23, 32), 1: (11, 52, 64), 2: (15, 112, 128), 3: (19, 236, 256)}[fmt]
    esz, bias, p = w // 8, (1 << (ew - 1)) - 1, mw + 1

    def enc(s, e, m):
        return ((s << (w - 1)) | (e << mw) | m).to_bytes(esz, "little")

    def frac():
        return int.from_bytes(rng.bytes(esz), "little") % (1 << mw)

    def sub():
        m = frac() >> int(rng.integers(0, mw))
        return m or 1

    aa, bb = [], []
    for i in range(n):
        sa, sb = (int(x) for x in rng.integers(0, 2, size=2))
        if kind == "mul":
            # lsb exponent of b = biased - bias - mw
            e = bias + mw + (-2 * p - 1) + int(rng.integers(-3, 4))
            x, y = enc(sa, 0, sub()), enc(sb, e, frac())
            if i % 2:
                x, y = y, x
        elif i % 10 == 0:
            x, y = enc(sa, 0, sub()), enc(sb, 0, sub())
        elif i % 2:
            x, y = enc(sa, 0, sub()), enc(sb, bias + mw + int(rng.integers(-3, 4)), frac())
        else:
            x, y = enc(sa, bias + int(rng.integers(-1, 5)), frac()), enc(sb, 0, sub())
        aa.append(x)
        bb.append(y)
    return (np.frombuffer(b"".join(aa), dtype=np.uint8).copy(),
            np.frombuffer(b"".join(bb), dtype=np.uint8).copy())


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_oracle_matches_reference_library(oracle, ref, fmt):
    """Every elementwise opcode the engine runs, all modes, 600 cases each."""
    rng = np.random.default_rng(7 + fmt)
    dev = ctypes.c_void_p()
    assert ref.cft_open(None, 0, ctypes.byref(dev)) == 0
    esz = {0: 4, 1: 8, 2: 16, 3: 32}[fmt]
    n = 600
    for op in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20,
               21, 22, 23, 26, 27, 30, 40, 256, 257):
        for rnd in range(5):
            a, b, c = (_rand_enc(rng, fmt, n) for _ in range(3))
            d = np.empty(n * esz, dtype=np.uint8)
            f = ctypes.c_uint32()
            bus = ctypes.c_uint32()
            if op == 256:
                st = ref.cft_div(dev, fmt, rnd, a.ctypes.data, b.ctypes.data, d.ctypes.data,
                                 n, ctypes.byref(f), ctypes.byref(bus))
            elif op == 257:
                st = ref.cft_sqrt(dev, fmt, rnd, a.ctypes.data, d.ctypes.data, n,
                                  ctypes.byref(f), ctypes.byref(bus))
            else:
                st = ref.cft_run(dev, op, fmt, rnd, a.ctypes.data, b.ctypes.data,
                                 c.ctypes.data, d.ctypes.data, n, ctypes.byref(f),
                                 ctypes.byref(bus))
            assert st == 0
            o, fl = oracle_elem(oracle, fmt, op, rnd, a, b, c)
            assert np.array_equal(o, d), (fmt, op, rnd)
            assert fl == f.value, (fmt, op, rnd)


def _fp256(sign, ef, frac):
    return ((sign << 255) | (ef << 236) | frac).to_bytes(32, "little")


def _one(oracle, op, rnd, a, b, c, fmt=3):
    arr = lambda x: np.frombuffer(x, dtype=np.uint8).copy() if x is not None else None
    d, fl = oracle_elem(oracle, fmt, op, rnd, arr(a), arr(b), arr(c), threads=1)
    return int.from_bytes(d.tobytes(), "little"), fl


def test_closed_forms(oracle):
    """Closed-form answers a plausible rounding mistake would break."""
    B = 262143
    one = _fp256(0, B, 0)
    zero = _fp256(0, 0, 0)
    # x*1 + 0 == x exactly, for a normal and a subnormal x
    for x in (_fp256(1, B + 77, 0x123456789 << 100), _fp256(0, 0, 12345)):
        r, fl = _one(oracle, 0, 0, x, one, zero)
        assert r == int.from_bytes(x, "little") and fl == 0
    # (1 + 2^-236)^2 = 1 + 2^-235 + 2^-472: RNE/RTZ/RDN drop the tail, RUP adds an ulp
    x = _fp256(0, B, 1)
    for rnd, want in ((0, 2), (1, 2), (2, 2), (3, 3), (4, 2)):
        r, fl = _one(oracle, 3, rnd, x, x, None)
        assert r == int.from_bytes(_fp256(0, B, want), "little"), rnd
        assert fl == 16
    # an exact tie: 1 + 2^-237 (= 1 + half an ulp): RNE to 1, RMM away, RUP away
    half = _fp256(0, B - 237, 0)
    for rnd, want in ((0, 0), (1, 0), (2, 0), (3, 1), (4, 1)):
        r, fl = _one(oracle, 1, rnd, one, None, half)
        assert r == int.from_bytes(_fp256(0, B, want), "little"), rnd
    # overflow: max * 2 -> inf in RNE, max in RTZ, with OF|NX
    mx = _fp256(0, 2 * B, (1 << 236) - 1)
    two = _fp256(0, B + 1, 0)
    r, fl = _one(oracle, 3, 0, mx, two, None)
    assert r == int.from_bytes(_fp256(0, 2 * B + 1, 0), "little") and fl == 4 | 16
    r, fl = _one(oracle, 3, 1, mx, two, None)
    assert r == int.from_bytes(mx, "little") and fl == 4 | 16
    # tininess after rounding: s = the largest subnormal; s * (1 + 2^-236) =
    # 2^emin - 2^(emin-472).  RNE rounds it to 2^emin, and rounded at full
    # precision it is 2^emin too: not tiny, so INEXACT alone.  RTZ keeps s:
    # tiny and inexact, so UNDERFLOW as well.
    s = ((1 << 236) - 1).to_bytes(32, "little")
    r, fl = _one(oracle, 3, 0, s, _fp256(0, B, 1), None)
    assert r == int.from_bytes(_fp256(0, 1, 0), "little") and fl == 16
    r, fl = _one(oracle, 3, 1, s, _fp256(0, B, 1), None)
    assert r == int.from_bytes(s, "little") and fl == 16 | 8
    # exact cancellation is +0 except under RDN
    x = _fp256(0, B + 3, 999)
    for rnd in range(5):
        r, fl = _one(oracle, 2, rnd, x, None, x)
        assert r == (1 << 255 if rnd == 2 else 0) and fl == 0
    # 0 * inf is invalid; qNaN + 0*inf is a quiet NaN with no invalid
    inf = _fp256(0, 2 * B + 1, 0)
    qnan = _fp256(0, 2 * B + 1, 1 << 235)
    r, fl = _one(oracle, 0, 0, zero, inf, one)
    assert fl == 1
    r, fl = _one(oracle, 0, 0, zero, inf, qnan)
    assert fl == 0 and r == int.from_bytes(qnan, "little")


def test_div_sqrt_closed_forms(oracle):
    B = 262143
    for k in (2, 3, 5, 7, 9, 1 << 20):
        # sqrt(k^2) == k exactly; (k*k)/k == k exactly
        kk = k * k
        ek = k.bit_length() - 1
        enc = lambda v, e: _fp256(0, B + e, (v << (236 - e)) & ((1 << 236) - 1))
        r, fl = _one(oracle, 257, 0, enc(kk, kk.bit_length() - 1), None, None)
        assert r == int.from_bytes(enc(k, ek), "little") and fl == 0
        r, fl = _one(oracle, 256, 0, enc(kk, kk.bit_length() - 1), enc(k, ek), None)
        assert r == int.from_bytes(enc(k, ek), "little") and fl == 0
    # 1/3 in RNE: 0.0101..01|01.. rounds down; RUP rounds up; flags NX
    one = _fp256(0, B, 0)
    three = _fp256(0, B + 1, 1 << 235)
    rd, _ = _one(oracle, 256, 1, one, three, None)
    ru, fl = _one(oracle, 256, 3, one, three, None)
    assert ru == rd + 1 and fl == 16
    # brute force against Python's exact integers on small significands
    rng = np.random.default_rng(3)
    for _ in range(200):
        x = int(rng.integers(1, 1 << 30)) | 1
        y = int(rng.integers(1, 1 << 30)) | 1
        ex, ey = x.bit_length() - 1, y.bit_length() - 1
        a = _fp256(0, B + ex, (x << (236 - ex)) & ((1 << 236) - 1))
        b = _fp256(0, B + ey, (y << (236 - ey)) & ((1 << 236) - 1))
        r, fl = _one(oracle, 256, 1, a, b, None)          # RTZ quotient
        # exact: q = floor(x * 2^s / y) with s chosen for 237 bits
        s = 237 + ey - ex + 2
        q = (x << s) // y
        L = q.bit_length()
        q >>= (L - 237)
        got_frac = r & ((1 << 236) - 1)
        assert got_frac == q & ((1 << 236) - 1)
