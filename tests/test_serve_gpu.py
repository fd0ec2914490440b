"""The serving loop on the GPU: a short mixed request stream (elementwise
binary128/binary256 FMA/DIV/SQRT on resident windows, Lorenz-63, Lorenz-96
and Henon-Heiles ensembles) served with batching, every job equal to the
same job run alone; the same stream served one job a launch gives the same
bits."""

import numpy as np
import pytest

from cftamd import CftError, _lib
from cftamd import serve as S
from conftest import oracle_elem
from seqlib import oracle_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deep(deep_device):
    return deep_device


def test_mixed_stream_each_job_as_alone(deep):
    reqs = S.requests(seed=7, rate=400, duration=0.25, pool_n=200_000, mix=(0.7, 0.1, 0.1, 0.1))
    pools = S.Pools(deep, reqs, seed=1, pool_n=200_000, verify_every=3)
    for r in reqs:                      # short ODE jobs for a test
        if r["kind"] >= 3:
            r["steps"] = np.maximum(1, np.asarray(r["steps"]) // 200).astype(np.uint32)
    srv = S.Server(deep, workers=4)
    res, wall = S.serve(srv, pools, reqs)
    stats = srv.stats()
    srv.close()
    summ = S.summarize(reqs, res, wall)
    assert summ["statuses_ok"]
    v = S.verify(deep, pools, reqs, res)
    assert v["mismatches"] == 0 and v["checked"] == len(reqs)
    assert stats["jobs"] == len(reqs)
    assert stats["max_batch_jobs"] >= 1 and stats["slices"] >= 1
    pools.free()


def mixed_requests(seed, pool_n, per=2, duration=0.2):
    """Every job kind in both elementwise formats, `per` of each, and every
    ODE (Lorenz-96 at a small and a large ring), arriving uniformly over
    `duration` seconds in a seeded order."""
    rng = np.random.default_rng(seed)
    reqs = []
    for kind in (0, 1, 2):
        for fmt in (2, 3):
            for _ in range(per):
                n = int(np.exp(rng.uniform(np.log(1e3), np.log(pool_n))))
                reqs.append({"kind": kind, "fmt": fmt, "n": n,
                             "off": int(rng.integers(0, pool_n - n + 1))})
    for _ in range(per):
        reqs.append({"kind": 3, "seed": int(rng.integers(1 << 31))})
        reqs.append({"kind": 5, "seed": int(rng.integers(1 << 31))})
    reqs.append({"kind": 4, "seed": int(rng.integers(1 << 31)), "k": 8})
    reqs.append({"kind": 4, "seed": int(rng.integers(1 << 31)), "k": 256})
    order = rng.permutation(len(reqs))
    times = np.sort(rng.uniform(0, duration, len(reqs)))
    return [dict(reqs[i], t=float(t)) for i, t in zip(order, times)]


def test_every_served_job_against_the_oracle(deep, oracle):
    """Each served job's output bits and flags against the oracle on the
    same operands: elementwise FMA / DIV / SQRT in binary128 and binary256
    (each job its own output window, raw device memory), and the final
    states, flags and STATUS
    of the Lorenz-63, Lorenz-96 and Henon-Heiles jobs against the oracle's
    chain of segments.  Programs run shortened (steps / 200) so the oracle
    finishes in seconds."""
    pool_n = 20_000
    reqs = mixed_requests(11, pool_n)
    pools = S.Pools(deep, reqs, seed=5, pool_n=pool_n, device_out=True)
    for r in reqs:
        if r["kind"] >= 3:
            r["steps"] = np.maximum(1, np.asarray(r["steps"]) // 200).astype(np.uint32)
    srv = S.Server(deep, workers=4)
    res, _wall = S.serve(srv, pools, reqs)
    srv.close()
    assert all(r.status == 0 for r in res)
    pools.states.from_device()
    kinds = set()
    for r, got in zip(reqs, res):
        if r["kind"] <= 2:
            fmt, esz, n = r["fmt"], (16 if r["fmt"] == 2 else 32), r["n"]
            win = slice(r["off"] * esz, (r["off"] + n) * esz)
            a, b, c, sa = (x.host[win] for x in pools.ops[fmt])
            if r["kind"] == 0:
                want, wfl = oracle_elem(oracle, fmt, 0, 0, a, b, c)
            elif r["kind"] == 1:
                want, wfl = oracle_elem(oracle, fmt, 256, 0, a, b, None)
            else:
                want, wfl = oracle_elem(oracle, fmt, 257, 0, sa, None, None)
            out = pools.served(r)
            assert np.array_equal(out, want), (r, np.flatnonzero(out != want)[:4])
            assert got.flags == wfl, (r, got.flags, wfl)
            kinds.add((r["kind"], fmt))
        else:
            lo, hi = r["state_off"], r["state_off"] + r["state_bytes"]
            want, wfl, wst = oracle_batch(oracle, pools.image_of(r), pools.bank_of(r),
                                          pools.state0[lo:hi], np.asarray(r["steps"]), 3,
                                          depth=2048)
            assert np.array_equal(pools.states.host[lo:hi], want), r["kind"]
            assert (got.flags, got.bus) == (wfl, wst), (r["kind"], got.flags, wfl, wst)
            kinds.add((r["kind"], 3))
    assert kinds == {(k, f) for k in (0, 1, 2) for f in (2, 3)} | {(3, 3), (4, 3), (5, 3)}
    pools.free()


def test_short_slices_give_the_same_bits(deep):
    # slices of a fraction of a millisecond: every program job runs in many
    # slices, joined and left by others as they arrive and finish
    reqs = S.requests(seed=9, rate=300, duration=0.2, pool_n=50_000, mix=(0.2, 0.3, 0.2, 0.3))
    pools = S.Pools(deep, reqs, seed=3, pool_n=50_000, verify_every=1)
    for r in reqs:
        if r["kind"] >= 3:
            r["steps"] = np.maximum(1, np.asarray(r["steps"]) // 100).astype(np.uint32)
    srv = S.Server(deep, workers=2)
    srv.set_slice(0.05)
    res, _wall = S.serve(srv, pools, reqs)
    stats = srv.stats()
    srv.close()
    n_prog = sum(r["kind"] >= 3 for r in reqs)
    # a slice advances a lane by one segment at this slice time (Lorenz-63
    # and Henon-Heiles segments of 100 trips, Lorenz-96 of 20), so the
    # longest job alone spans this many slices
    seg = {3: 100, 4: 20, 5: 100}
    most = max(-(-int(np.asarray(r["steps"]).max()) // seg[r["kind"]]) for r in reqs
               if r["kind"] >= 3)
    assert most > 1 and stats["slices"] >= most
    assert all(r.status == 0 for r in res)
    v = S.verify(deep, pools, reqs, res)
    assert v["mismatches"] == 0 and v["programs_alone"] == n_prog
    pools.free()


def test_unbatched_gives_the_same_bits(deep):
    reqs = S.requests(seed=8, rate=200, duration=0.2, pool_n=100_000, mix=(0.6, 0.2, 0.1, 0.1))
    pools = S.Pools(deep, reqs, seed=2, pool_n=100_000, verify_every=1)
    for r in reqs:
        if r["kind"] >= 3:
            r["steps"] = np.maximum(1, np.asarray(r["steps"]) // 500).astype(np.uint32)
    srv = S.Server(deep, workers=1)
    srv.set_batching(1)
    res, _wall = S.serve(srv, pools, reqs)
    assert srv.stats()["max_batch_jobs"] == 1
    srv.close()
    v = S.verify(deep, pools, reqs, res)
    assert v["mismatches"] == 0
    pools.free()


def test_impure_job_past_the_batch_cap_runs(deep, oracle):
    """A job of a program the wide runner does not take (a quiet region is
    flag control) with more lanes than a whole batch gathers (2^22) is still
    admitted, alone, and runs to its end: every lane's x = 1.5 becomes
    1.5 * 1.5 + 0 = 2.25, exactly."""
    from cftamd import Program
    from cftamd import programs as P
    img = P.image(3, [P.repeat(1), P.ldl(3, 0), P.quiet(), P.alu(0, 3, 3, 3, 0, 0),
                      P.endquiet(), P.stl(3, 0), P.endrep(), P.halt()],
                  scratch_io=1 | (1 << 16))
    prog = Program(deep, img)
    n = (1 << 22) + 1
    st = deep.alloc(n * 32)
    st.host.reshape(n, 32)[:] = np.frombuffer(P.dec(3, "1.5").to_bytes(32, "little"), np.uint8)
    st.to_device()
    srv = S.Server(deep, workers=1)
    t = srv.submit_program(prog, st.data_ptr, n * 32, np.ones(n, np.uint32))
    r = srv.wait(t)
    srv.close()
    assert r.status == 0 and r.flags == 0 and r.bus == 0
    st.from_device()
    want = np.frombuffer(P.dec(3, "2.25").to_bytes(32, "little"), np.uint8)
    assert (st.host.reshape(n, 32) == want).all()
    prog.free()
    st.free()


def test_submit_rules(deep):
    srv = S.Server(deep, workers=1)
    host = np.zeros(64, np.uint8)           # not a resident buffer
    with pytest.raises(CftError) as ei:
        srv.submit_elem(_lib.JOB_ELEM, 0, 3, 0, host.ctypes.data, host.ctypes.data,
                        host.ctypes.data, host.ctypes.data, 2)
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    buf = deep.alloc(64 * 32)
    buf.to_device()
    with pytest.raises(CftError):
        srv.submit_elem(_lib.JOB_ELEM, 24, 3, 0, buf.data_ptr, buf.data_ptr, buf.data_ptr,
                        buf.data_ptr, 64)          # a reduction opcode
    with pytest.raises(CftError):
        srv.set_slice(-1.0)
    t = srv.submit_elem(_lib.JOB_ELEM, 1, 3, 0, buf.data_ptr, None, buf.data_ptr,
                        buf.data_ptr, 64)
    assert srv.wait(t).status == 0
    srv.close()
    buf.free()
