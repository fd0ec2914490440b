"""The sequencer on the GPU: libcft's interpreter (k_seq.hip) through
cft_program_run_ex against the certificates' boundary states, and against
the oracle's executor on the same inputs - ODE programs over many lanes and
random programs over every control code and format."""

import ctypes
import hashlib
import struct

import numpy as np
import pytest

from cftamd import Device, Program, _lib, CftError
from cftamd import programs as P
from seqlib import ESZ, random_case, rand_values, run_gpu, run_oracle, diff
from test_seq import FMT, cert_bank, cert_cases, cert_run, certs, widen64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,run", cert_cases())
def test_certificates_on_gpu(device, name, run):
    assert cert_run(None, name, run, lambda case: run_gpu(device, case)) is None


@pytest.mark.parametrize("name", sorted(certs()))
def test_digest_matches_certificate(device, name):
    c = certs()[name]
    fmt = FMT[c["format"]]
    prog = Program(device, P.GENERATORS[c["system"]](fmt))
    for r in c["runs"]:
        bank = cert_bank(name, c["system"], fmt, r["index"])
        assert prog.digest(bank).hex() == r["program_digest"]
    prog.free()


def ode_case(rng, sysname, fmt, lanes, trips, k=None):
    kw = {"k": k} if k else {}
    img = P.with_trips(P.GENERATORS[sysname](fmt, **kw), trips)
    slots = {"lorenz63-rk4": 3, "henonheiles-lf": 4}.get(sysname, k)
    bank = np.frombuffer(P.bank_bytes(sysname, fmt), dtype=np.uint8).copy()
    vals = 1 + rng.uniform(-1e-3, 1e-3, size=lanes * slots)
    if sysname == "henonheiles-lf":
        vals = rng.uniform(-0.3, 0.3, size=lanes * slots)
    sin = np.frombuffer(struct.pack(f"<{vals.size}d", *vals), dtype=np.uint8).copy()
    if fmt == 3:
        sin = widen64(sin)
    return {"image": img, "bank": bank, "a": np.zeros(lanes * ESZ[fmt], dtype=np.uint8),
            "b": None, "c": None, "n": lanes, "fmt": fmt, "scratch_in": sin,
            "n_out": slots, "max_dep": 0, "mask": None}


@pytest.mark.parametrize("fmt", [1, 3])
def test_odes_many_lanes_match_oracle(device, oracle, fmt):
    rng = np.random.default_rng(10 + fmt)
    for sysname, lanes, trips, k in (("lorenz63-rk4", 1000, 20, None),
                                     ("henonheiles-lf", 777, 30, None),
                                     ("lorenz96-rk4", 130, 3, 16),
                                     ("lorenz96-rk4", 64, 2, 40)):
        case = ode_case(rng, sysname, fmt, lanes, trips, k)
        print(f"ode {sysname} fmt {fmt} lanes {lanes} trips {trips} k {k}", flush=True)
        assert diff(run_gpu(device, case), run_oracle(oracle, case, threads=16)) == [], sysname


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_random_programs_match_oracle(device, oracle, fmt):
    rng = np.random.default_rng(100 + fmt)
    for it in range(60):
        n = int(rng.choice([1, 7, 64, 65, 200, 513]))
        case = random_case(rng, fmt, n, budget=int(rng.integers(5, 120)))
        print(f"random program fmt {fmt} case {it} n {n}", flush=True)
        assert diff(run_gpu(device, case), run_oracle(oracle, case)) == [], it


def test_index_tables_gather(device, oracle):
    """A run with index tables equals the dense run on the gathered inputs
    (host/src/seq_exec.h:689-711: CFT_IDX_NONE reads +0)."""
    rng = np.random.default_rng(5)
    fmt, n = 3, 300
    esz = ESZ[fmt]
    for it in range(8):
        case = random_case(rng, fmt, n, budget=40)
        if case["scratch_in"] is None:
            continue
        prog = Program(device, case["image"])
        n_in = case["scratch_in"].nbytes // (n * esz)
        pool_n = 97
        srcs = {k: rand_values(rng, fmt, pool_n) for k in "abc"}
        spool = rand_values(rng, fmt, pool_n)
        idx = {k: rng.integers(0, pool_n, size=n).astype(np.uint32) for k in "abc"}
        isi = rng.integers(0, pool_n, size=n * n_in).astype(np.uint32)
        for t in list(idx.values()) + [isi]:
            t[rng.random(t.size) < 0.1] = 0xFFFFFFFF

        def gather(pool, t):
            out = np.zeros(t.size * esz, dtype=np.uint8)
            for i, j in enumerate(t):
                if j != 0xFFFFFFFF:
                    out[i * esz:(i + 1) * esz] = pool[j * esz:(j + 1) * esz]
            return out
        dense = dict(case)
        dense["a"], dense["b"], dense["c"] = (gather(srcs[k], idx[k]) for k in "abc")
        dense["scratch_in"] = gather(spool, isi)
        want = run_oracle(oracle, dense)
        got = run_gpu(device, dense, prog)
        assert diff(got, want) == []
        # the same through the tables
        from seqlib import _outputs
        out = _outputs(case)
        fl, st = prog.run(n, srcs["a"], srcs["b"], srcs["c"], bank=case["bank"],
                          scratch_in=spool, scratch_out=out["scratch_out"],
                          deposits=out["deposits"], counts=out["counts"],
                          lane_mask=case["mask"], lane_flags=out["lane_flags"],
                          idx_a=idx["a"], idx_b=idx["b"], idx_c=idx["c"], idx_a_src=pool_n,
                          idx_b_src=pool_n, idx_c_src=pool_n, idx_scratch_in=isi,
                          idx_scratch_src=pool_n)
        out["flags"], out["status"] = fl, st
        assert diff(out, want) == []
        prog.free()


def test_segments_chain_in_resident_buffers(device, oracle):
    """Lorenz-63 state kept in one resident buffer, scratch-in and
    scratch-out the same bytes, five segments in a row: the same as the
    oracle's five segments."""
    rng = np.random.default_rng(3)
    lanes = 4096
    case = ode_case(rng, "lorenz63-rk4", 3, lanes, 10)
    prog = Program(device, case["image"])
    st = device.alloc(case["scratch_in"].nbytes)
    st.host[:] = case["scratch_in"]
    st.to_device()
    bank = device.alloc(case["bank"].nbytes)
    bank.host[:] = case["bank"]
    bank.to_device()
    a = device.alloc(case["a"].nbytes)
    a.host[:] = 0
    a.to_device()
    want = case["scratch_in"].copy()
    for seg in range(5):
        fl, status = prog.run(lanes, a, bank=bank, scratch_in=st, scratch_out=st)
        c2 = dict(case)
        c2["scratch_in"] = want
        o = run_oracle(oracle, c2, threads=16)
        want = o["scratch_out"]
        assert fl == o["flags"] and status == o["status"] == 0
    st.from_device()
    assert np.array_equal(st.host, want)
    assert st.info().resident_binds >= 10
    for b in (st, bank, a):
        b.free()
    prog.free()


def test_l96_ring_256_at_depth_2048(oracle):
    """Lorenz-96 at K = 256 needs 1280 scratch slots: a handle opened at
    depth 2048 runs it; one at the default depth refuses it at load."""
    rng = np.random.default_rng(8)
    case = ode_case(rng, "lorenz96-rk4", 3, 70, 1, k=256)
    with Device("hip:0") as d256:
        with pytest.raises(CftError) as ei:
            Program(d256, case["image"])
        assert ei.value.status == _lib.ERR_UNSUPPORTED
    with Device("hip:0", scratch_depth=2048) as d:
        got = run_gpu(d, case)
    assert diff(got, run_oracle(oracle, case, depth=2048, threads=16)) == []


def test_run_rules(device):
    img = P.lorenz63(3)
    prog = Program(device, img)
    lanes = 5
    bank = np.frombuffer(P.bank_bytes("lorenz63-rk4", 3), dtype=np.uint8).copy()
    a = np.zeros(lanes * 32, dtype=np.uint8)
    sin = np.zeros(lanes * 3 * 32, dtype=np.uint8)
    sout = np.zeros_like(sin)

    def st(**kw):
        args = dict(n=lanes, a=a, bank=bank, scratch_in=sin, scratch_out=sout)
        args.update(kw)
        try:
            prog.run(**args)
            return 0
        except CftError as e:
            return e.status
    assert st() == 0
    # n = 0 is a run of no lanes: OK with empty blocks, and a block sized
    # for other lanes is refused (program.c:1685-1700 sizes it n x slots)
    assert st(n=0, scratch_in=None, scratch_out=None) == 0
    assert st(n=0) == _lib.ERR_INVALID_ARGUMENT
    assert st(a=None) == _lib.ERR_INVALID_ARGUMENT
    assert st(bank=None) == _lib.ERR_INVALID_ARGUMENT
    assert st(bank=bank[:-32]) == _lib.ERR_INVALID_ARGUMENT
    assert st(scratch_in=sin[:-32]) == _lib.ERR_INVALID_ARGUMENT
    assert st(scratch_out=None) == _lib.ERR_INVALID_ARGUMENT
    assert st(lane_mask=np.zeros(2, dtype=np.uint8)) == _lib.ERR_INVALID_ARGUMENT
    assert st(lane_flags=np.zeros(4, dtype=np.uint8)) == _lib.ERR_INVALID_ARGUMENT
    assert st(idx_a=np.zeros(lanes, dtype=np.uint32)) == _lib.ERR_INVALID_ARGUMENT
    assert st(idx_a=np.full(lanes, 3, dtype=np.uint32), idx_a_src=3) == \
        _lib.ERR_INVALID_ARGUMENT
    assert st(idx_a_src=4) == _lib.ERR_INVALID_ARGUMENT
    # the positional entry points refuse what they cannot carry
    L = _lib.lib()
    assert L.cft_program_run(prog.handle, a.ctypes.data, None, None, None, None, lanes,
                             None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.cft_program_run_bank(prog.handle, bank.ctypes.data, bank.nbytes, a.ctypes.data,
                                  None, None, None, None, lanes, None, None) == \
        _lib.ERR_INVALID_ARGUMENT
    bad = _lib.CftRunArgs()
    bad.struct_size = ctypes.sizeof(bad) - 8
    assert L.cft_program_run_ex(prog.handle, ctypes.byref(bad)) == _lib.ERR_INVALID_ARGUMENT
    info = prog.info()
    assert (info.n_insns, info.n_consts, info.flags, info.n_scratch_in, info.n_scratch_out,
            info.scratch_used) == (62, 7, 3, 3, 3, 3)
    prog.free()


def test_plain_program_with_own_constants(device, oracle):
    """An image that carries its constants, run through cft_program_run
    with deposits: r0..r2 from the streams."""
    fmt = 2
    rng = np.random.default_rng(4)
    consts = [P.dec(fmt, "1.5"), P.dec(fmt, "-0.25")]
    insns = [P.alu(P.FMA, 3, 0, 1, P.K(0)), P.deposit(3), P.alu(P.MUL, 4, 3, P.K(1)),
             P.deposit(4), P.alu(P.ADD, 5, 4, 0, 2), P.deposit(5), P.halt()]
    img = P.image(fmt, insns, consts=consts, max_deposits=3)
    n = 1000
    case = {"image": img, "bank": None, "a": rand_values(rng, fmt, n),
            "b": rand_values(rng, fmt, n), "c": rand_values(rng, fmt, n), "n": n,
            "fmt": fmt, "scratch_in": None, "n_out": 0, "max_dep": 3, "mask": None}
    want = run_oracle(oracle, case)
    prog = Program(device, img)
    dep = np.zeros(n * 3 * 16, dtype=np.uint8)
    cnt = np.zeros(n, dtype=np.uint32)
    fl, bus = ctypes.c_uint32(), ctypes.c_uint32()
    st = _lib.lib().cft_program_run(prog.handle, case["a"].ctypes.data, case["b"].ctypes.data,
                                    case["c"].ctypes.data, dep.ctypes.data,
                                    cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n,
                                    ctypes.byref(fl), ctypes.byref(bus))
    assert st == 0
    assert np.array_equal(dep, want["deposits"]) and np.array_equal(cnt, want["counts"])
    assert fl.value == want["flags"] and bus.value == want["status"]
    prog.free()
