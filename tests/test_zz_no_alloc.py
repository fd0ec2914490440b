"""cftx_reduce_noalloc, the reduction that allocates nothing, through the C
ABI against cft_reduce on the same operands: operands in device memory
(hipMalloc through the HIP runtime libcft itself uses), two calls in a row
on every format and reduction opcode, the second smaller than the first,
results and flags bit for bit; the empty array's identity; the
infinity-ahead-of-NaN rule of sumSquare and sumAbs on operands holding both;
and the refusals (too many elements, host memory, misalignment) made before
anything runs."""

import ctypes

import numpy as np
import pytest

from cftamd import _lib
from cftamd._lib import lib

pytestmark = pytest.mark.gpu

ESZ = {0: 4, 1: 8, 2: 16, 3: 32}
EW = {0: 8, 1: 11, 2: 15, 3: 19}
OPS = (24, 25, 28, 29, 31)
MAX_N = {f: (1 << 27) // ESZ[f] for f in ESZ}


def values(rng, fmt, n, spread=6):
    """n finite binary{32..256} values within `spread` binades of 1 (either
    sign), a tenth of them up to 40 binades away; dense little-endian bytes."""
    nl, ew = ESZ[fmt] // 4, EW[fmt]
    bias = (1 << (ew - 1)) - 1
    x = rng.integers(0, 1 << 32, size=(n, nl), dtype=np.uint64).astype(np.uint32)
    wide = rng.random(n) < 0.1
    e = bias + np.where(wide, rng.integers(-40, 41, n), rng.integers(-spread, spread + 1, n))
    s = rng.integers(0, 2, n).astype(np.uint64)
    fr = x[:, nl - 1].astype(np.uint64) & ((1 << (31 - ew)) - 1)
    x[:, nl - 1] = ((s << 31) | (e.astype(np.uint64) << (31 - ew)) | fr).astype(np.uint32)
    return x.reshape(-1).view(np.uint8).copy()


def special(fmt, kind, sign=0):
    nl, ew = ESZ[fmt] // 4, EW[fmt]
    emask = ((1 << ew) - 1) << (31 - ew)
    x = np.zeros(nl, np.uint32)
    top = (sign << 31) | emask
    if kind == "qnan":
        top |= 1 << (30 - ew)
    elif kind == "snan":
        x[0] = 1
    x[nl - 1] = top
    return x.view(np.uint8)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libcft is linked
    against (the same loaded library, so the same device address space)."""

    def __init__(self):
        lib()                                   # libcft (and its runtime) first
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        V = ctypes.c_void_p
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [V]
        self.rt.hipMemcpy.argtypes = [V, V, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipMemGetAddressRange.argtypes = [ctypes.POINTER(V),
                                                  ctypes.POINTER(ctypes.c_size_t), V]
        self.live = []

    def upload(self, host):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(1, host.nbytes)) == 0
        self.live.append(p.value)
        assert self.rt.hipMemcpy(p.value, host.ctypes.data, host.nbytes, 1) == 0
        return p.value

    def empty(self, nbytes, fill=0xA5):
        return self.upload(np.full(nbytes, fill, np.uint8))

    def download(self, ptr, nbytes):
        out = np.empty(nbytes, np.uint8)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def hip(device):
    h = Hip()
    yield h
    h.free_all()


def noalloc(dev, op, fmt, rnd, a, b, n, d):
    fl = ctypes.c_uint32(0xDEAD)
    st = lib().cftx_reduce_noalloc(dev.handle, op, fmt, rnd, a, b, d, n, ctypes.byref(fl),
                                   None)
    return st, fl.value


def compare(hip, device, fmt, op, rnd, a, b, n):
    want, wfl = device.reduce(op, fmt, rnd, a, b if op == 25 else None, n=n)
    da, db, dd = hip.upload(a), hip.upload(b), hip.empty(ESZ[fmt])
    st, fl = noalloc(device, op, fmt, rnd, da, db, n, dd)
    assert st == 0, lib().cft_last_error()
    got = hip.download(dd, ESZ[fmt])
    hip.free_all()
    assert got.tobytes() == want.tobytes(), (fmt, op, rnd, n)
    assert fl == wfl, (fmt, op, rnd, n, fl, wfl)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_two_calls_match_cft_reduce(hip, device, fmt):
    rng = np.random.default_rng(4100 + fmt)
    big = {0: MAX_N[0], 1: (1 << 22) + 7, 2: 300_001, 3: MAX_N[3]}[fmt]
    small = 1000
    a, b = values(rng, fmt, big), values(rng, fmt, big)
    for op in OPS:
        rnd = int(rng.integers(0, 5))
        compare(hip, device, fmt, op, rnd, a, b, big)
        compare(hip, device, fmt, op, rnd, a[: small * ESZ[fmt]].copy(),
                b[: small * ESZ[fmt]].copy(), small)


def test_sizes_through_the_tree(hip, device):
    rng = np.random.default_rng(4200)
    for n in (1, 2, 15, 16, 17, 255, 256, 257, 4097, 65537):
        a, b = values(rng, 3, n), values(rng, 3, n)
        for op in OPS:
            compare(hip, device, 3, op, int(rng.integers(0, 5)), a, b, n)


def test_specials_and_the_94_rule(hip, device):
    rng = np.random.default_rng(4300)
    for fmt in (1, 3):
        esz = ESZ[fmt]
        for mix in (("inf", "snan"), ("inf", "qnan"), ("snan",), ("inf",), ("inf", "ninf")):
            n = 333
            a, b = values(rng, fmt, n), values(rng, fmt, n)
            for k, kind in enumerate(mix):
                pos = int(rng.integers(0, n))
                sp = special(fmt, "inf", 1) if kind == "ninf" else special(fmt, kind)
                a[pos * esz:(pos + 1) * esz] = sp
            for op in OPS:
                compare(hip, device, fmt, op, 0, a, b, n)
    # a second call with no specials after one with them: nothing carries
    a, b = values(rng, 3, 50), values(rng, 3, 50)
    for op in OPS:
        compare(hip, device, 3, op, 0, a, b, 50)


def test_empty_array_identity(hip, device):
    for fmt in ESZ:
        for op in OPS:
            want, wfl = device.reduce(op, fmt, 0, None, None, n=0)
            dd = hip.empty(ESZ[fmt])
            st, fl = noalloc(device, op, fmt, 0, None, None, 0, dd)
            assert st == 0 and fl == wfl == 0
            assert hip.download(dd, ESZ[fmt]).tobytes() == want.tobytes(), (fmt, op)
            hip.free_all()


def test_refusals(hip, device):
    t, d = hip.empty(4096, 0), hip.empty(32, 0)
    # more elements than the scratch array's partials hold (refused before
    # any operand is looked at)
    st, _ = noalloc(device, 24, 3, 0, t, None, MAX_N[3] + 1, d)
    assert st == _lib.ERR_INVALID_ARGUMENT
    assert b"scratch" in lib().cft_last_error()
    # host memory as an operand or as the result
    host = np.zeros(4096, np.uint8)
    st, _ = noalloc(device, 24, 3, 0, host.ctypes.data, None, 64, d)
    assert st == _lib.ERR_INVALID_ARGUMENT
    st, _ = noalloc(device, 24, 3, 0, t, None, 64, host.ctypes.data)
    assert st == _lib.ERR_INVALID_ARGUMENT
    # a window running one element past the end of its allocation, as the
    # runtime reports the allocation's range
    base, size = ctypes.c_void_p(), ctypes.c_size_t()
    assert hip.rt.hipMemGetAddressRange(ctypes.byref(base), ctypes.byref(size),
                                        ctypes.c_void_p(t)) == 0
    past = (base.value + size.value - t) // 32 + 1
    st, _ = noalloc(device, 24, 3, 0, t, None, past, d)
    assert st == _lib.ERR_INVALID_ARGUMENT
    # misaligned
    st, _ = noalloc(device, 24, 3, 0, t + 4, None, 64, d)
    assert st == _lib.ERR_INVALID_ARGUMENT
    hip.free_all()
    # the handle still works afterwards
    a = values(np.random.default_rng(1), 3, 100)
    compare(hip, device, 3, 24, 0, a, a, 100)
