"""The batched runner (cftx_program_run_batch) on the GPU: the persistent
form and the per-op form against the oracle's chain of segment runs, over
ODE ensembles of uneven length; launch shapes and graph replay change no
bit."""

import numpy as np
import pytest

from cftamd import CftError, Device, Program, _lib, odes
from cftamd import programs as P
from seqlib import oracle_batch

pytestmark = pytest.mark.gpu


def run(dev, img, bank, state, steps, **kw):
    prog = Program(dev, img)
    st = state.copy()
    fl, status = prog.run_batch(st, steps, bank=bank, **kw)
    prog.free()
    return st, fl, status


@pytest.mark.parametrize("fmt", [1, 3])
def test_lorenz63_uneven_matches_oracle(device, oracle, fmt):
    img, bank, state, steps = odes.lorenz63(300, seed=fmt, fmt=fmt, steps_lo=1, steps_hi=450)
    steps[:7] = [0, 1, 99, 100, 101, 200, 450]
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, fmt)
    lf = np.full(steps.size, 0xEE, dtype=np.uint8)
    got, fl, st = run(device, img, bank, state, steps, lane_flags=lf)
    assert np.array_equal(got, want) and (fl, st) == (wfl, wst)
    assert set(lf.tolist()) <= {0, 16} and lf[0] == 0 and lf[1] == 16


@pytest.mark.parametrize("mode,lpw,wpg,graph", [
    (_lib.BATCH_PERSISTENT, 1, 1, False), (_lib.BATCH_PERSISTENT, 7, 2, False),
    (_lib.BATCH_PERSISTENT, 64, 4, True), (_lib.BATCH_PERSISTENT, 0, 0, True),
    (_lib.BATCH_PER_OP, 0, 0, False)])
def test_shapes_and_forms_agree(device, oracle, mode, lpw, wpg, graph):
    img, bank, state, steps = odes.lorenz63(500, seed=9, fmt=3, steps_lo=1, steps_hi=300)
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 3)
    prog = Program(device, img)
    for rep in range(2 if graph else 1):        # a graph replay the second time
        st = state.copy()
        fl, status = prog.run_batch(st, steps, bank=bank, mode=mode, lanes_per_wave=lpw,
                                    waves_per_group=wpg, graph=graph)
        assert np.array_equal(st, want) and (fl, status) == (wfl, wst), rep
    launches = _lib.lib().cftx_batch_launches(device.handle)
    if mode == _lib.BATCH_PERSISTENT:
        assert launches == 2
    else:
        assert launches > 53 * int(steps.max())
    prog.free()


def test_lorenz96_and_henonheiles(device, oracle):
    for k, img, bank, ids, state, steps in odes.lorenz96(40, seed=2, fmt=1, steps_lo=1,
                                                         steps_hi=60, ks=[8, 13, 24]):
        want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 1)
        got, fl, st = run(device, img, bank, state, steps)
        assert np.array_equal(got, want) and (fl, st) == (wfl, wst), k
        got2, fl2, st2 = run(device, img, bank, state, steps, mode=_lib.BATCH_PER_OP)
        assert np.array_equal(got2, want) and (fl2, st2) == (wfl, wst), k
    img, bank, state, steps = odes.henonheiles(100, seed=4, fmt=3, steps_lo=1, steps_hi=350)
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 3)
    got, fl, st = run(device, img, bank, state, steps)
    assert np.array_equal(got, want) and (fl, st) == (wfl, wst)


def test_state_in_resident_buffer(device, oracle):
    img, bank, state, steps = odes.lorenz63(1000, seed=5, fmt=3, steps_lo=50, steps_hi=250)
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 3)
    prog = Program(device, img)
    buf = device.alloc(state.nbytes)
    buf.host[:] = state
    buf.to_device()
    fl, st = prog.run_batch(buf, steps, bank=bank, graph=True)
    buf.from_device()
    assert np.array_equal(buf.host, want) and (fl, st) == (wfl, wst)
    buf.free()
    prog.free()


def test_batch_rules(device):
    img, bank, state, steps = odes.lorenz63(10, seed=1, fmt=3, steps_lo=1, steps_hi=9)
    prog = Program(device, img)

    def status(**kw):
        args = dict(state=state.copy(), steps=steps, bank=bank)
        args.update(kw)
        try:
            prog.run_batch(**args)
            return 0
        except CftError as e:
            return e.status
    assert status() == 0
    assert status(state=state[:-32].copy()) == _lib.ERR_INVALID_ARGUMENT
    assert status(bank=None) == _lib.ERR_INVALID_ARGUMENT
    assert status(lanes_per_wave=65) == _lib.ERR_INVALID_ARGUMENT
    assert status(waves_per_group=3) == _lib.ERR_INVALID_ARGUMENT
    assert status(mode=2) == _lib.ERR_INVALID_ARGUMENT
    assert status(mode=_lib.BATCH_PER_OP, lane_flags=np.zeros(10, np.uint8)) == \
        _lib.ERR_INVALID_ARGUMENT
    prog.free()
    # programs that are not a segment: deposits, no state, no loop
    for bad in (P.image(3, [P.repeat(2), P.endrep(), P.halt()], max_deposits=1,
                        scratch_in=1, scratch_out=1),
                P.image(3, [P.repeat(2), P.endrep(), P.halt()]),
                P.image(3, [P.halt()], scratch_in=1, scratch_out=1)):
        p = Program(device, bad)
        with pytest.raises(CftError) as ei:
            p.run_batch(np.zeros(32, np.uint8), np.ones(1, np.uint32))
        assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
        p.free()


@pytest.mark.parametrize("system,fmt,k,widths", [
    ("lorenz63-rk4", 3, None, [1, 2, 4, 8]), ("lorenz63-rk4", 1, None, [4]),
    ("henonheiles-lf", 3, None, [2, 4]), ("lorenz96-rk4", 1, 12, [2, 16, 64]),
    ("lorenz96-rk4", 3, 40, [8, 64])])
def test_wide_widths_match_oracle(device, oracle, system, fmt, k, widths):
    rng = np.random.default_rng(31)
    if system == "lorenz63-rk4":
        img, bank, state, steps = odes.lorenz63(200, seed=3, fmt=fmt, steps_lo=1, steps_hi=250)
    elif system == "henonheiles-lf":
        img, bank, state, steps = odes.henonheiles(150, seed=3, fmt=fmt, steps_lo=1,
                                                   steps_hi=250)
    else:
        (_k, img, bank, _ids, state, steps), = odes.lorenz96(50, seed=3, fmt=fmt, steps_lo=1,
                                                             steps_hi=45, ks=[k])
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, fmt)
    prog = Program(device, img)
    for w in widths:
        for graph in (False, True):
            st = state.copy()
            lf = np.full(steps.size, 0xEE, dtype=np.uint8)
            fl, status = prog.run_batch(st, steps, bank=bank, lanes_per_instance=w,
                                        graph=graph, lane_flags=lf)
            assert _lib.lib().cftx_batch_width(device.handle) == w
            assert np.array_equal(st, want) and (fl, status) == (wfl, wst), (w, graph)
            assert set(lf.tolist()) <= {0, 16}
    st = state.copy()
    fl, status = prog.run_batch(st, steps, bank=bank)              # the width it chooses
    assert np.array_equal(st, want) and (fl, status) == (wfl, wst)
    prog.free()


def test_wide_random_pure_programs(device, oracle):
    from test_wide import pure_program
    from seqlib import rand_values, ESZ
    rng = np.random.default_rng(91)
    for it in range(24):
        fmt = int(rng.integers(0, 4))
        n_io = int(rng.integers(1, 6))
        img, consts = pure_program(rng, fmt, n_io, int(rng.integers(0, 20)))
        n = int(rng.integers(1, 300))
        state = rand_values(rng, fmt, n * n_io)
        steps = rng.integers(0, 7, size=n).astype(np.uint32)
        want, wfl, wst = oracle_batch(oracle, img, None, state, steps, fmt)
        prog = Program(device, img)
        for w in (1, 2, 4, 16, 64):
            st = state.copy()
            fl, status = prog.run_batch(st, steps, lanes_per_instance=w)
            assert np.array_equal(st, want) and (fl, status) == (wfl, wst), (it, fmt, w)
        prog.free()


def test_wide_refuses_impure_programs(device):
    img = P.image(3, [P.repeat(2), P.stx(3, 4), P.endrep(), P.halt()], scratch_in=1,
                  scratch_out=1)
    prog = Program(device, img)
    with pytest.raises(CftError) as ei:
        prog.run_batch(np.zeros(32, np.uint8), np.ones(1, np.uint32), lanes_per_instance=4)
    assert ei.value.status == _lib.ERR_UNSUPPORTED
    st = np.zeros(32, np.uint8)
    prog.run_batch(st, np.ones(1, np.uint32))                     # narrow: fine
    prog.free()
