"""GPU parity: the HIP kernels through the C ABI against the oracle and the
reference-written golden vectors, bit for bit and flag for flag.

Flags are checked per element without one call per element: the golden
vectors carry each element's flags, the elements are grouped by that value,
and each group is run as one call whose OR must be exactly the group's value.
"""

import ctypes

import numpy as np
import pytest

from conftest import oracle_elem

pytestmark = pytest.mark.gpu

OPS = {"fma": 0, "add": 1, "sub": 2, "mul": 3, "div": 256, "sqrt": 257}


def gpu(device, fmt, op, rnd, a, b, c):
    if op == 256:
        return device.div(fmt, rnd, a, b)
    if op == 257:
        return device.sqrt(fmt, rnd, a)
    return device.run(op, fmt, rnd, a, b, c)


def operands(opn, g):
    b = None if opn in ("add", "sub", "sqrt") else g["b"]
    c = g["c"] if opn in ("fma", "add", "sub") else None
    return g["a"], b, c


@pytest.mark.parametrize("name,fmt", [("fp128", 2), ("fp256", 3)])
@pytest.mark.parametrize("opn", list(OPS))
def test_golden_bits_and_flags(device, golden, name, fmt, opn):
    g = golden[name]
    esz = 16 if fmt == 2 else 32
    a, b, c = operands(opn, g)
    for rnd in range(5):
        d, fl = gpu(device, fmt, OPS[opn], rnd, a, b, c)
        want = g[f"{opn}_r{rnd}"]
        bad = np.nonzero((d.reshape(-1, esz) != want.reshape(-1, esz)).any(axis=1))[0]
        assert bad.size == 0, f"{name} {opn} rnd {rnd}: elements {bad[:8]} differ"
        wf = g[f"{opn}_r{rnd}_flags"]
        assert fl == int(np.bitwise_or.reduce(wf))
        for v in np.unique(wf):
            idx = np.nonzero(wf == v)[0]
            sel = lambda x: None if x is None else x.reshape(-1, esz)[idx].reshape(-1).copy()
            _, gfl = gpu(device, fmt, OPS[opn], rnd, sel(a), sel(b), sel(c))
            assert gfl == int(v), f"{name} {opn} rnd {rnd}: flag group {v} gave {gfl}"


def _rand(rng, fmt, n):
    from test_oracle import _rand_enc
    return _rand_enc(rng, fmt, n)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_every_opcode_every_format(device, oracle, fmt):
    rng = np.random.default_rng(500 + fmt)
    for op in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20,
               21, 22, 23, 26, 27, 30, 40, 200, 256, 257):
        for rnd in range(5):
            a, b, c = (_rand(rng, fmt, 700) for _ in range(3))
            want, wf = oracle_elem(oracle, fmt, op, rnd, a, b, c)
            got, gf = gpu(device, fmt, op, rnd, a, b, c)
            assert np.array_equal(got, want), (fmt, op, rnd)
            assert gf == wf, (fmt, op, rnd)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_division_and_root_next_to_one(device, oracle, fmt):
    """Quotients within a few ulps of 1 (quotient digits B-1, where the
    digit estimate is capped) and the square roots of both operands."""
    from test_oracle import _next_to
    a, b = _next_to(np.random.default_rng(950 + fmt), fmt, 3000)
    for x, y in ((a, b), (b, a)):
        for rnd in range(5):
            for op, yy in ((256, y), (257, None)):
                want, wf = oracle_elem(oracle, fmt, op, rnd, x, yy, None)
                got, gf = gpu(device, fmt, op, rnd, x, yy, None)
                assert np.array_equal(got, want) and gf == wf, (fmt, op, rnd)


def test_flags_do_not_carry_between_calls(device, oracle):
    """A call's kernel ORs its flags into the handle's flag slots, which the
    call's gather leaves zero: calls of 2^16 + 3 elements (257 workgroups,
    every slot) and of 5 (one workgroup) alternate between operands that
    raise flags and exact ones (1 * 1 + 0, 1 / 1), each call's result and
    flags the oracle's."""
    rng = np.random.default_rng(77)
    one = (262143 << 236).to_bytes(32, "little")          # binary256 1.0

    def exact(n):
        x = np.frombuffer(one * n, dtype=np.uint8).copy()
        return x, x.copy(), np.zeros(n * 32, dtype=np.uint8)

    def raising(n):
        return tuple(_rand(rng, 3, n) for _ in range(3))

    big, small = (1 << 16) + 3, 5
    for kind, n in (("raising", big), ("exact", small), ("raising", small), ("exact", big),
                    ("raising", big), ("exact", big), ("exact", small)):
        a, b, c = (raising if kind == "raising" else exact)(n)
        for op in (0, 256):
            cc = c if op == 0 else None
            want, wf = oracle_elem(oracle, 3, op, 0, a, b, cc)
            got, gf = gpu(device, 3, op, 0, a, b, cc)
            assert np.array_equal(got, want) and gf == wf, (kind, n, op, gf, wf)
            assert kind == "raising" or gf == 0, (kind, n, op, gf)


@pytest.mark.parametrize("cfg,fmt", [(1, 2), (2, 3)])
def test_workload_configs_all_ops(device, oracle, cfg, fmt):
    """Configs 1 and 2 (near-boundary FMA cases included) at 2^16."""
    from cftamd import workload
    a, b, c = workload.elementwise(cfg, 1 << 16, seed=0)
    absa = a.copy()
    esz = 16 if fmt == 2 else 32
    absa.reshape(-1, esz)[:, esz - 1] &= 0x7F
    for opn, op in OPS.items():
        for rnd in range(5):
            x = absa if opn == "sqrt" else a
            bb = None if opn in ("add", "sub", "sqrt") else b
            cc = c if opn in ("fma", "add", "sub") else None
            want, wf = oracle_elem(oracle, fmt, op, rnd, x, bb, cc)
            got, gf = gpu(device, fmt, op, rnd, x, bb, cc)
            assert np.array_equal(got, want), (cfg, opn, rnd)
            assert gf == wf, (cfg, opn, rnd)


def test_full_size_fma_config2(device, oracle):
    """Config 2's binary256 FMA at its full N = 2^24, RNE and RDN, every bit."""
    from cftamd import workload
    n = 1 << 24
    a, b, c = workload.elementwise(2, n, seed=0)
    for rnd in (0, 2):
        got, gf = device.run(0, 3, rnd, a, b, c)
        want, wf = oracle_elem(oracle, 3, 0, rnd, a, b, c, threads=16)
        assert gf == wf
        assert np.array_equal(got, want)


def test_call_rules_on_device(device):
    from cftamd import _lib
    lib = _lib.lib()
    a = np.zeros(32 * 4, dtype=np.uint8)
    fl = ctypes.c_uint32(77)
    # n == 0: OK, no flags
    assert lib.cft_run(device.handle, 0, 3, 0, None, None, None, None, 0, ctypes.byref(fl),
                       None) == 0 and fl.value == 0
    # a NULL operand the opcode reads is refused; one it does not read is +0
    d = np.empty_like(a)
    assert lib.cft_run(device.handle, 0, 3, 0, a.ctypes.data, None, a.ctypes.data,
                       d.ctypes.data, 4, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_run(device.handle, 1, 3, 0, a.ctypes.data, None, a.ctypes.data,
                       d.ctypes.data, 4, None, None) == 0
    # reductions go through cft_reduce, not here
    assert lib.cft_run(device.handle, 24, 3, 0, a.ctypes.data, None, None, d.ctypes.data, 4,
                       None, None) == _lib.ERR_INVALID_ARGUMENT
    # bad rounding attribute, bad format, opcode past a byte
    assert lib.cft_run(device.handle, 0, 3, 5, a.ctypes.data, a.ctypes.data, a.ctypes.data,
                       d.ctypes.data, 4, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_run(device.handle, 0, 4, 0, a.ctypes.data, a.ctypes.data, a.ctypes.data,
                       d.ctypes.data, 4, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_run(device.handle, 256, 3, 0, a.ctypes.data, a.ctypes.data,
                       a.ctypes.data, d.ctypes.data, 4, None, None) == \
        _lib.ERR_INVALID_ARGUMENT


def test_aliasing_scalar_and_sticky(device, oracle):
    rng = np.random.default_rng(9)
    a, b, c = (_rand(rng, 3, 1000) for _ in range(3))
    want, wf = oracle_elem(oracle, 3, 0, 0, a, b, c)
    x = a.copy()
    _, fl = device.run(0, 3, 0, x, b, c, out=x)          # d aliases a
    assert np.array_equal(x, want) and fl == wf
    # scalar operand: a[0] applies to every element
    from cftamd import _lib
    lib = _lib.lib()
    d = np.empty_like(a)
    f = ctypes.c_uint32()
    args = _lib.CftElemArgs()
    args.struct_size = ctypes.sizeof(args)
    args.a, args.b, args.c, args.d = a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data
    args.n = 1000
    args.scalar_mask = 1
    args.flags_out = ctypes.pointer(f)
    assert lib.cft_run_ex(device.handle, 0, 3, 0, ctypes.byref(args)) == 0
    aa = np.tile(a[:32], 1000)
    want2, wf2 = oracle_elem(oracle, 3, 0, 0, aa, b, c)
    assert np.array_equal(d, want2) and f.value == wf2
    # index tables: a gathered through idx_a (CFT_IDX_NONE reads +0)
    idx = (np.arange(1000, dtype=np.uint32) * 7) % 1000
    idx[::13] = 0xFFFFFFFF
    args.scalar_mask = 0
    args.idx_a = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    args.idx_a_src = 1000
    assert lib.cft_run_ex(device.handle, 0, 3, 0, ctypes.byref(args)) == 0
    ga = np.zeros_like(a).reshape(-1, 32)
    ok = idx != 0xFFFFFFFF
    ga[ok] = a.reshape(-1, 32)[idx[ok]]
    want3, wf3 = oracle_elem(oracle, 3, 0, 0, ga.reshape(-1), b, c)
    assert np.array_equal(d, want3) and f.value == wf3
    idx[5] = 1000                                        # past the source: refused
    assert lib.cft_run_ex(device.handle, 0, 3, 0, ctypes.byref(args)) == \
        _lib.ERR_INVALID_ARGUMENT
    # the sticky word ORs calls' flags and is lowered only on request
    device.lower_flags()
    device.run(0, 3, 0, a, b, c)
    assert device.save_all_flags() == wf
    device.lower_flags(16)
    assert device.save_all_flags() == wf & ~16


def test_resident_buffers(device, oracle):
    from cftamd import workload
    n = 1 << 14
    a, b, c = workload.elementwise(2, n, seed=3)
    bufs = [device.alloc(n * 32) for _ in range(4)]
    for buf, src in zip(bufs, (a, b, c)):
        buf.host[:] = src
        buf.to_device()
    d = bufs[3]
    _, fl = device.run(0, 3, 0, bufs[0], bufs[1], bufs[2], out=d, n=n)
    info = d.info()
    assert info.device_authority == 1 and bufs[0].info().resident_binds == 1
    # to_device is refused while a run's results are unread
    with pytest.raises(Exception):
        d.to_device()
    d.from_device()
    want, wf = oracle_elem(oracle, 3, 0, 0, a, b, c)
    assert np.array_equal(d.host, want) and fl == wf
    # a buffer written on the host after to_device is restaged on its next bind
    bufs[0].host[:] = c
    bufs[0].to_device()
    device.run(0, 3, 0, bufs[0], bufs[1], bufs[2], out=d, n=n)
    d.from_device()
    want2, _ = oracle_elem(oracle, 3, 0, 0, c, b, c)
    assert np.array_equal(d.host, want2)
    for buf in bufs:
        buf.free()


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_exponent_settled_quotients_and_products(device, oracle, fmt):
    """Division and multiplication at the thresholds where a result is
    settled by its exponent alone, both sides of them: bits against the
    oracle, and flags per exponent class of the operands."""
    from test_oracle import _edges
    rng = np.random.default_rng(700 + fmt)
    esz = 4 << fmt
    for kind, op in (("div", 256), ("mul", 3)):
        a, b = _edges(rng, fmt, 4000, kind)
        for rnd in range(5):
            want, wf = oracle_elem(oracle, fmt, op, rnd, a, b, None)
            d, fl = gpu(device, fmt, op, rnd, a, b, None)
            bad = np.nonzero((d.reshape(-1, esz) != want.reshape(-1, esz)).any(axis=1))[0]
            assert bad.size == 0, f"fmt {fmt} {kind} rnd {rnd}: elements {bad[:8]} differ"
            assert fl == wf, (fmt, kind, rnd)
