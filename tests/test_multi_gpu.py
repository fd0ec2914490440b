"""Several programs in one launch (cftx_program_run_multi) on the GPU: every
part's states, flags and lane flags equal the oracle's chain of segment
runs of that part alone, whatever the planner or the caller picks for lanes
an instance; the refusals."""

import numpy as np
import pytest

from cftamd import CftError, Program, _lib, odes, run_multi, tune_multi
from cftamd import programs as P
from seqlib import ESZ, oracle_batch, rand_values
from test_wide import pure_program

pytestmark = pytest.mark.gpu


def test_ode_mix_matches_the_oracle(device, oracle):
    parts, want = [], []
    img, bank, state, steps = odes.lorenz63(200, seed=1, fmt=3, steps_lo=1, steps_hi=300)
    steps[:5] = [0, 1, 99, 100, 101]
    parts.append(dict(img=img, bank=bank, state=state, steps=steps))
    img, bank, state, steps = odes.henonheiles(100, seed=2, fmt=3, steps_lo=1, steps_hi=350)
    parts.append(dict(img=img, bank=bank, state=state, steps=steps))
    for k, img, bank, _ids, state, steps in odes.lorenz96(12, seed=3, fmt=3, steps_lo=1,
                                                          steps_hi=30, ks=[8, 40]):
        parts.append(dict(img=img, bank=bank, state=state, steps=steps))
    progs = []
    for p in parts:
        want.append(oracle_batch(oracle, p["img"], p["bank"], p["state"], p["steps"], 3))
        progs.append(Program(device, p["img"]))
        p["prog"] = progs[-1]
        p["state"] = p["state"].copy()
        p["lane_flags"] = np.full(p["steps"].size, 0xEE, np.uint8)
    fl, status, widths = run_multi(device, parts)
    wfl = wst = 0
    for p, (ws, f, s) in zip(parts, want):
        assert np.array_equal(p["state"], ws)
        wfl |= f
        wst |= s
        assert set(p["lane_flags"].tolist()) <= {0, 16}
        assert (p["lane_flags"][p["steps"] == 0] == 0).all()
    assert (fl, status) == (wfl, wst)
    assert all(w in (1, 2, 4, 8, 16, 32, 64) for w in widths)
    # lane flags are those of the part run alone
    for p in parts:
        st = p["state"].copy()
        lf = np.zeros(p["steps"].size, np.uint8)
        p["prog"].run_batch(st, p["steps"], bank=p["bank"], lane_flags=lf)
        assert np.array_equal(lf, p["lane_flags"])
    for q in progs:
        q.free()


def test_every_width_in_one_launch(device, oracle):
    # one program in seven parts, each at its own lanes an instance
    img, bank, state, steps = odes.lorenz63(70, seed=5, fmt=1, steps_lo=1, steps_hi=250)
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 1)
    prog = Program(device, img)
    esz = ESZ[1] * 3
    parts = []
    for k in range(7):
        sl = slice(10 * k, 10 * k + 10)
        parts.append(dict(prog=prog, bank=bank, state=state[sl.start * esz:sl.stop * esz].copy(),
                          steps=steps[sl], lanes_per_instance=1 << k))
    fl, status, widths = run_multi(device, parts)
    assert widths == [1, 2, 4, 8, 16, 32, 64]
    got = np.concatenate([p["state"] for p in parts])
    assert np.array_equal(got, want) and (fl, status) == (wfl, wst)
    prog.free()


def test_lorenz96_across_k_matches_run_batch(deep_device, oracle):
    device = deep_device
    groups = odes.lorenz96(96, seed=7, fmt=1, steps_lo=1, steps_hi=80)
    parts, progs = [], []
    for k, img, bank, _ids, state, steps in groups:
        progs.append(Program(device, img))
        parts.append(dict(prog=progs[-1], bank=bank, state=state.copy(), steps=steps, k=k,
                          img=img, state0=state))
    fl, status, widths = run_multi(device, parts)
    ks = [p["k"] for p in parts]
    assert 256 in ks or max(ks) >= 200
    for p in parts:
        st = p["state0"].copy()
        p["prog"].run_batch(st, p["steps"], bank=p["bank"])
        assert np.array_equal(st, p["state"]), p["k"]
    for p in parts[:: max(1, len(parts) // 4)]:      # a few against the oracle too
        ws, _f, _s = oracle_batch(oracle, p["img"], p["bank"], p["state0"], p["steps"], 1,
                                  depth=2048)
        assert np.array_equal(ws, p["state"]), p["k"]
    for q in progs:
        q.free()


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_random_pure_programs_match_the_oracle(device, oracle, fmt):
    rng = np.random.default_rng(100 + fmt)
    parts, want, progs = [], [], []
    for _ in range(6):
        n_io = int(rng.integers(1, 6))
        img, _consts = pure_program(rng, fmt, n_io, int(rng.integers(0, 20)))
        n = int(rng.integers(1, 90))
        state = rand_values(rng, fmt, n * n_io)
        steps = rng.integers(0, 6, size=n).astype(np.uint32)
        want.append(oracle_batch(oracle, img, None, state, steps, fmt))
        progs.append(Program(device, img))
        parts.append(dict(prog=progs[-1], state=state.copy(), steps=steps))
    fl, status, _w = run_multi(device, parts)
    wfl = 0
    for p, (ws, f, _s) in zip(parts, want):
        assert np.array_equal(p["state"], ws)
        wfl |= f
    assert fl == wfl
    for q in progs:
        q.free()


def test_state_in_resident_buffer(device, oracle):
    img, bank, state, steps = odes.henonheiles(300, seed=8, fmt=3, steps_lo=1, steps_hi=200)
    want, wfl, wst = oracle_batch(oracle, img, bank, state, steps, 3)
    prog = Program(device, img)
    buf = device.alloc(state.nbytes)
    buf.host[:] = state
    buf.to_device()
    fl, status, _w = run_multi(device, [dict(prog=prog, bank=bank, state=buf, steps=steps)])
    buf.from_device()
    assert np.array_equal(buf.host, want) and (fl, status) == (wfl, wst)
    buf.free()
    prog.free()


def test_refusals(device):
    img1, bank1, state1, steps1 = odes.lorenz63(4, seed=1, fmt=1, steps_lo=1, steps_hi=5)
    img3, bank3, state3, steps3 = odes.lorenz63(4, seed=1, fmt=3, steps_lo=1, steps_hi=5)
    p1, p3 = Program(device, img1), Program(device, img3)
    with pytest.raises(CftError) as ei:                     # two formats
        run_multi(device, [dict(prog=p1, bank=bank1, state=state1.copy(), steps=steps1),
                           dict(prog=p3, bank=bank3, state=state3.copy(), steps=steps3)])
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(CftError) as ei:                     # state size
        run_multi(device, [dict(prog=p3, bank=bank3, state=state3[:-32].copy(), steps=steps3)])
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(CftError) as ei:                     # lanes an instance
        run_multi(device, [dict(prog=p3, bank=bank3, state=state3.copy(), steps=steps3,
                                lanes_per_instance=3)])
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    impure = Program(device, P.image(3, [P.repeat(2), P.setact(1), P.endrep(), P.halt()],
                                     scratch_in=1, scratch_out=1))
    with pytest.raises(CftError) as ei:
        run_multi(device, [dict(prog=impure, state=np.zeros(32, np.uint8), steps=[1])])
    assert ei.value.status == _lib.ERR_UNSUPPORTED
    assert p3.prepare(0x7F) & 0x7F
    with pytest.raises(CftError):
        impure.prepare(1)
    # no lanes with trips: nothing runs, states untouched
    st = state3.copy()
    fl, status, _w = run_multi(device, [dict(prog=p3, bank=bank3, state=st,
                                             steps=np.zeros(4, np.uint32))])
    assert np.array_equal(st, state3) and (fl, status) == (0, 0)
    for q in (p1, p3, impure):
        q.free()


def test_planner_options_change_no_bit(deep_device):
    device = deep_device
    groups = odes.lorenz96(64, seed=11, fmt=3, steps_lo=1, steps_hi=40, ks=[8, 48, 136, 256])
    progs = [Program(device, g[1]) for g in groups]
    img, bank, state, steps = odes.lorenz63(100, seed=12, fmt=3, steps_lo=1, steps_hi=250)
    progs.append(Program(device, img))

    def parts():
        out = [dict(prog=p, bank=g[2], state=g[4].copy(), steps=g[5])
               for p, g in zip(progs, groups)]
        out.append(dict(prog=progs[-1], bank=bank, state=state.copy(), steps=steps))
        return out

    base = parts()
    f0, s0, _w = run_multi(device, base)
    best, trials = tune_multi(device, parts())
    assert trials and best in [t for _s, t in trials]
    for knobs in (dict(waves_per_group=1), dict(waves_per_group=2, alone=0.3),
                  dict(waves_per_cu=8), dict(waves_per_cu=2, alone=1.0), best):
        try:
            got = parts()
            f, s, _w = run_multi(device, got, **knobs)
        except CftError as e:            # a plan no program fits is refused, not run
            assert e.status == _lib.ERR_UNSUPPORTED, knobs
            continue
        assert (f, s) == (f0, s0)
        for a, b in zip(base, got):
            assert np.array_equal(a["state"], b["state"]), knobs
    with pytest.raises(CftError):
        run_multi(device, parts(), waves_per_cu=13)
    with pytest.raises(CftError):
        run_multi(device, parts(), alone=1.5)
    for q in progs:
        q.free()
