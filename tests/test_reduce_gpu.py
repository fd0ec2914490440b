"""GPU parity of cft_reduce and cft_reduce_seg: the HIP tree through the C ABI
against the oracle's reductions (oracle/rd_oracle.c, itself pinned to the
reference's library in test_reduce.py), bit for bit and flag for flag, over
every opcode, format and rounding attribute; segments of every shape; the
9.4 override per segment; resident operands and results; the argument rules;
and at full size the properties the contract fixes without an oracle run:
exact integer sums, the maximum, and the tree's own composition (a
power-of-two segment is a node of the whole array's tree, so reducing the
segments' results reproduces the whole-array call)."""

import ctypes
import struct

import numpy as np
import pytest

from cftamd import Buffer, CftError
from test_reduce import ESZ, OPS, red_values, run_oc

pytestmark = pytest.mark.gpu

V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)


@pytest.fixture(scope="module")
def oc(oracle):
    oracle.oc_reduce.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                                 ctypes.c_size_t, U32P]
    oracle.oc_reduce_seg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V,
                                     ctypes.c_size_t, ctypes.c_size_t, U32P]
    return oracle


def gpu(device, fmt, op, rnd, a, b, n, seg=None):
    b = b if op == 25 else None
    return device.reduce(op, fmt, rnd, a, b, n=n, seg=seg)


SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 17, 31, 63, 64, 65, 255, 256, 257, 1000, 1023,
         4097]


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_whole_array_matches_oracle(device, oc, fmt):
    rng = np.random.default_rng(100 + fmt)
    for op in OPS.values():
        for n in SIZES:
            rnd = int(rng.integers(0, 5))
            a = red_values(rng, fmt, n, special=0.002 if n > 64 else 0.03)
            b = red_values(rng, fmt, n, special=0.002 if n > 64 else 0.03)
            want = run_oc(oc, fmt, op, rnd, a, b, n)
            got = gpu(device, fmt, op, rnd, a, b, n)
            assert got[0].tobytes() == want[0].tobytes(), (op, n, rnd)
            assert got[1] == want[1], (op, n, rnd, got[1], want[1])


def test_flags_do_not_carry_between_calls(device, oc):
    """SUM, DOT and MAXALL take their flags through the handle's flag slots,
    zero between calls: calls of 2^18 + 5 elements (every slot) and of 7
    alternate between values whose sums round and ones (exact sums), each
    call's result and flags the oracle's."""
    rng = np.random.default_rng(78)
    one = (262143 << 236).to_bytes(32, "little")          # binary256 1.0
    big, small = (1 << 18) + 5, 7
    for kind, n in (("rounding", big), ("exact", small), ("rounding", small), ("exact", big),
                    ("rounding", big), ("exact", big)):
        if kind == "exact":
            a = np.frombuffer(one * n, dtype=np.uint8).copy()
            b = a.copy()
        else:
            a = red_values(rng, 3, n, special=0.0)
            b = red_values(rng, 3, n, special=0.0)
        for op in (24, 25, 31):
            want = run_oc(oc, 3, op, 0, a, b, n)
            got = gpu(device, 3, op, 0, a, b, n)
            assert got[0].tobytes() == want[0].tobytes(), (kind, n, op)
            assert got[1] == want[1], (kind, n, op, got[1], want[1])
            assert kind == "rounding" or got[1] == 0, (kind, n, op, got[1])


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_every_rounding_attribute(device, oc, fmt):
    # sums that round at nearly every node, so the attribute shows everywhere
    rng = np.random.default_rng(200 + fmt)
    n = 777
    for op in (24, 25, 28, 29):
        a = red_values(rng, fmt, n, special=0.0)
        b = red_values(rng, fmt, n, special=0.0)
        for rnd in range(5):
            want = run_oc(oc, fmt, op, rnd, a, b, n)
            got = gpu(device, fmt, op, rnd, a, b, n)
            assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], (op, rnd)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_segments_match_oracle(device, oc, fmt):
    rng = np.random.default_rng(300 + fmt)
    n = 1680
    for op in OPS.values():
        for seg in (1, 2, 3, 4, 5, 7, 8, 12, 16, 35, 48, 80, 112, 240, 560, 1680):
            rnd = int(rng.integers(0, 5))
            a = red_values(rng, fmt, n, special=0.01)
            b = red_values(rng, fmt, n, special=0.01)
            want = run_oc(oc, fmt, op, rnd, a, b, n, seg)
            got = gpu(device, fmt, op, rnd, a, b, n, seg)
            assert got[0].tobytes() == want[0].tobytes(), (op, seg, rnd)
            assert got[1] == want[1], (op, seg, rnd)


def test_special_mixes_per_segment(device, oc):
    # infinities, NaNs of both kinds and zeros of both signs, so the 9.4 row
    # holds for some segments and not for their neighbours
    fmt = 1
    pool = [struct.pack("<d", v) for v in (float("inf"), float("-inf"), 0.0, -0.0, 1.5, -2.25)]
    pool += [struct.pack("<Q", 0x7FF8000000000000), struct.pack("<Q", 0x7FF0000000000001),
             struct.pack("<Q", 0xFFF0000000000002)]
    rng = np.random.default_rng(9)
    for trial in range(40):
        seg = int(rng.integers(1, 7))
        n = seg * int(rng.integers(1, 40))
        pick = lambda: np.frombuffer(b"".join(pool[int(rng.integers(0, len(pool)))]
                                              for _ in range(n)), np.uint8).copy()
        a, b = pick(), pick()
        for op in OPS.values():
            rnd = int(rng.integers(0, 5))
            want = run_oc(oc, fmt, op, rnd, a, b, n, seg)
            got = gpu(device, fmt, op, rnd, a, b, n, seg)
            assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], \
                (trial, op, seg)
            want = run_oc(oc, fmt, op, rnd, a, b, n)
            got = gpu(device, fmt, op, rnd, a, b, n)
            assert got[0].tobytes() == want[0].tobytes() and got[1] == want[1], (trial, op)


def test_identities_and_lone_elements(device):
    for fmt in range(4):
        esz = ESZ[fmt]
        d, fl = device.reduce(24, fmt, 0, None, n=0)
        assert d.tobytes() == bytes(esz) and fl == 0
        d, fl = device.reduce(31, fmt, 0, None, n=0)
        inf = {0: 0xFF800000, 1: 0xFFF0000000000000}
        if fmt in inf:
            assert int.from_bytes(d.tobytes(), "little") == inf[fmt] and fl == 0
    snan = np.frombuffer(struct.pack("<Q", 0xFFF0000000000001), np.uint8).copy()
    for op in (24, 31):
        d, fl = device.reduce(op, 1, 0, snan, n=1)
        assert d.tobytes() == snan.tobytes() and fl == 0
    d, fl = device.reduce(29, 1, 0, snan, n=1)
    assert struct.unpack("<Q", d.tobytes())[0] == 0x7FF0000000000001 and fl == 0
    d, fl = device.reduce(28, 1, 0, snan, n=1)
    assert struct.unpack("<Q", d.tobytes())[0] == 0x7FF8000000000000 and fl == 1
    # a segmented call over nothing writes nothing
    d = np.full(8, 0xAB, np.uint8)
    _, fl = device.reduce(24, 1, 0, None, out=d, n=0, seg=4)
    assert (d == 0xAB).all() and fl == 0


def test_argument_rules(device):
    a = np.zeros(64, np.uint8)
    for kw, st in [(dict(op=0), 1), (dict(op=26), 1), (dict(fmt=4), 1), (dict(rnd=5), 1),
                   (dict(op=25, b=None), 1), (dict(seg=0), 1), (dict(seg=3), 1),
                   (dict(a=None), 1)]:
        args = dict(op=24, fmt=1, rnd=0, a=a, b=a, n=8, seg=None)
        args.update(kw)
        with pytest.raises(CftError) as e:
            device.reduce(args["op"], args["fmt"], args["rnd"], args["a"], args["b"],
                          n=args["n"], seg=args["seg"])
        assert e.value.status == st, kw
    from cftamd._lib import lib
    fl = ctypes.c_uint32()
    assert lib().cft_reduce(device.handle, 24, 1, 0, a.ctypes.data, None, None, 8,
                            ctypes.byref(fl), None) == 1
    for op in OPS.values():
        for fmt in range(4):
            assert lib().cft_supports(device.handle, op, fmt) == 1
    caps = device.caps()
    assert caps.seq_features & 0x1000            # CFT_FEAT_REDUCE_SEG


def test_resident_operands_and_results(device, oc):
    fmt, n, seg = 3, 4096, 64
    esz = ESZ[fmt]
    rng = np.random.default_rng(17)
    a = red_values(rng, fmt, n, special=0.005)
    b = red_values(rng, fmt, n, special=0.005)
    ba, bb = Buffer(device, n * esz), Buffer(device, n * esz)
    bd = Buffer(device, (n // seg + 8) * esz)
    ba.host[:] = a
    bb.host[:] = b
    bd.host[:] = 0x5A
    try:
        for op in OPS.values():
            for rnd in (0, 3):
                want = run_oc(oc, fmt, op, rnd, a, b, n, seg)
                # d at an offset inside a buffer: only its window changes
                out = bd.host[8 * esz:8 * esz + (n // seg) * esz]
                _, fl = device.reduce(op, fmt, rnd, ba, bb if op == 25 else None,
                                      out=out, n=n, seg=seg)
                bd.from_device()
                assert bd.host[8 * esz:].tobytes() == want[0].tobytes() and fl == want[1]
                assert (bd.host[:8 * esz] == 0x5A).all()
                want1 = run_oc(oc, fmt, op, rnd, a, b, n)
                got1 = device.reduce(op, fmt, rnd, ba, bb if op == 25 else None, n=n)
                assert got1[0].tobytes() == want1[0].tobytes() and got1[1] == want1[1]
        assert ba.info().resident_binds > 0
    finally:
        for x in (ba, bb, bd):
            x.free()


def test_exact_integer_sums_at_full_size(device):
    # integers below 2^20 in binary64: every partial of 2^24 of them is
    # exact, so the tree's sum is the integer sum in every attribute
    rng = np.random.default_rng(3)
    n = 1 << 24
    v = rng.integers(-(1 << 20), 1 << 20, n)
    a = v.astype(np.float64).view(np.uint8)
    for rnd in range(5):
        d, fl = device.reduce(24, 1, rnd, a, n=n)
        assert struct.unpack("<d", d.tobytes())[0] == float(v.sum()) and fl == 0
    w = v >> 6                                    # squares below 2^28, their sum below 2^52
    d, fl = device.reduce(28, 1, 0, w.astype(np.float64).view(np.uint8), n=n)
    assert struct.unpack("<d", d.tobytes())[0] == float((w * w).sum()) and fl == 0
    d, fl = device.reduce(29, 1, 0, a, n=n)
    assert struct.unpack("<d", d.tobytes())[0] == float(np.abs(v).sum())
    d, fl = device.reduce(24, 1, 0, a, n=n, seg=4096)
    want = v.reshape(-1, 4096).sum(axis=1).astype(np.float64)
    assert np.array_equal(d.view(np.float64), want) and fl == 0


def test_maximum_at_full_size(device):
    rng = np.random.default_rng(4)
    n = 10_000_019
    x = rng.standard_normal(n)
    d, fl = device.reduce(31, 1, 0, x.view(np.uint8), n=n)
    assert d.view(np.float64)[0] == x.max() and fl == 0
    d, fl = device.reduce(31, 1, 0, x.view(np.uint8)[:1000 * 9973 * 8], n=1000 * 9973,
                          seg=9973)
    assert np.array_equal(d.view(np.float64), x[:1000 * 9973].reshape(1000, -1).max(axis=1))


@pytest.mark.parametrize("fmt", [1, 3])
def test_power_of_two_segments_compose_to_the_whole(device, fmt):
    # T(0, n) over n = k * 2^j is the same tree over the k nodes of 2^j
    rng = np.random.default_rng(40 + fmt)
    n, seg = 1 << 20, 1 << 10
    a = red_values(rng, fmt, 4096, special=0.0)
    a = np.tile(a, n // 4096)
    a.reshape(n, -1)[:, -1] ^= rng.integers(0, 2, n, dtype=np.uint8) << 7    # signs
    for op, rnd in ((24, 0), (24, 2), (29, 3), (28, 4), (25, 1)):
        whole, fw = device.reduce(op, fmt, rnd, a, a[::-1].copy(), n=n)
        parts, fp = device.reduce(op, fmt, rnd, a, a[::-1].copy(), n=n, seg=seg)
        top, ft = device.reduce(24, fmt, rnd, parts, n=n // seg)
        assert whole.tobytes() == top.tobytes(), (op, rnd)
        assert fw == fp | ft, (op, rnd)


def test_results_over_their_own_operands(device, oc):
    # d aliasing a: the segments' results written over the array's own head,
    # from host memory and inside one resident buffer
    fmt, n, seg = 2, 3000, 30
    esz = ESZ[fmt]
    rng = np.random.default_rng(23)
    a = red_values(rng, fmt, n, special=0.01)
    for op in (24, 28, 31):
        want = run_oc(oc, fmt, op, 1, a, None, n, seg)
        x = a.copy()
        _, fl = device.reduce(op, fmt, 1, x, out=x[: (n // seg) * esz], n=n, seg=seg)
        assert x[: (n // seg) * esz].tobytes() == want[0].tobytes() and fl == want[1], op
        buf = Buffer(device, n * esz)
        try:
            buf.host[:] = a
            _, fl = device.reduce(op, fmt, 1, buf, out=buf.host[: (n // seg) * esz], n=n,
                                  seg=seg)
            buf.from_device()
            assert buf.host[: (n // seg) * esz].tobytes() == want[0].tobytes(), op
            assert buf.host[(n // seg) * esz:].tobytes() == a[(n // seg) * esz:].tobytes()
            assert fl == want[1], op
        finally:
            buf.free()
