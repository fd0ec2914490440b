"""The serving loop on the GPU: a short mixed request stream (elementwise
binary128/binary256 FMA/DIV/SQRT on resident windows, Lorenz-63, Lorenz-96
and Henon-Heiles ensembles) served with batching, every job equal to the
same job run alone; the same stream served one job a launch gives the same
bits."""

import numpy as np
import pytest

from cftamd import CftError, _lib
from cftamd import serve as S

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
