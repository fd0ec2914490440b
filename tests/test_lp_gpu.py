"""The eight-lane binary256 FMA (cft-fp256_amd/csrc/fx_lp.hpp, run by
lib/libcftprobe.so's cftp_lp_run: one operation over an aligned group of 8
lanes, a limb a lane) on the differential stream's hard cases: FMA, ADD,
SUB and MUL (steered to the FMA they are defined as), all five modes, every
element's bits and its own flags against the oracle, the fast path and the
hand-back to the one-lane arithmetic both exercised."""

import ctypes

import numpy as np
import pytest

import diffstream as ds

pytestmark = pytest.mark.gpu

V = ctypes.c_void_p


def lp_run(P, op, rnd, a, b, c):
    from cftamd import DeviceArray
    n, es = a.shape
    rt = DeviceArray.runtime()
    keep = []

    def up(x):
        if x is None:
            return None
        d = DeviceArray(x.nbytes)
        assert rt.hipMemcpy(d.ptr, x.ctypes.data, x.nbytes, 1) == 0
        keep.append(d)
        return d.ptr

    pa, pb, pc = up(a), up(b), up(c)
    dd, df, dfast = DeviceArray(n * es), DeviceArray(n), DeviceArray(n)
    try:
        assert P.cftp_lp_run(op, rnd, pa, pb, pc, dd.ptr, n, df.ptr, dfast.ptr) == 0
        return dd.download().reshape(n, es), df.download(), dfast.download()
    finally:
        for x in keep + [dd, df, dfast]:
            x.free()


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_eight_lane_fma_matches_the_oracle(device, op):
    P = ds.load_probe()
    P.cftp_lp_run.argtypes = [ctypes.c_int, ctypes.c_int, V, V, V, V, ctypes.c_size_t, V, V]
    O = ds.load_oracle()
    for rnd in range(5):
        a, b, c, fam = ds.generate(P, 3, op, 500 + rnd, 0, 14 * 4000)
        got, fl, fast = lp_run(P, op, rnd, a, b, c)
        want, _, lanes = ds.oracle_call(O, 3, op, rnd, a, b, c, ds.cpu_threads())
        bad = np.flatnonzero((got != want).any(axis=1) | (fl != lanes))
        assert bad.size == 0, (op, rnd, int(bad[0]), int(fam[bad[0]]))
        # both paths ran: the common families in the eight-lane path, the
        # specials handed back
        assert fast[fam == 7].mean() > 0.9 and fast[fam == 8].mean() < 0.5


def test_eight_lane_fma_chain(device):
    from cftamd import programs as Pg
    P = ds.load_probe()
    P.cftp_lp_chain_latency.argtypes = [ctypes.c_int, V, V, ctypes.POINTER(ctypes.c_double)]
    words = np.frombuffer(b"".join(int(v).to_bytes(32, "little") for v in
                                   (Pg.dec(3, "1.5"), Pg.dec(3, "1.0000001"), Pg.dec(3, "1e-30"))),
                          np.uint32).copy()
    out = np.zeros(10, np.uint32)
    us = ctypes.c_double()
    assert P.cftp_lp_chain_latency(64, words.ctypes.data, out.ctypes.data, ctypes.byref(us)) == 0
    assert out[9] == 0                     # every step on the fast path
    # the chain's value: 64 FMAs of the one-lane path on the oracle
    O = ds.load_oracle()
    x = words[:8].view(np.uint8).reshape(1, 32).copy()
    a = words[8:16].view(np.uint8).reshape(1, 32).copy()
    c = words[16:24].view(np.uint8).reshape(1, 32).copy()
    fl = 0
    for _ in range(64):
        x, f, _l = ds.oracle_call(O, 3, 0, 0, x, a, c, 1)
        fl |= f
    assert (out[:8].view(np.uint8) == x.reshape(-1)).all() and out[8] == fl
