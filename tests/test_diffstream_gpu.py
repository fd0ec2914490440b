"""The counter-based differential stream on the GPU (tests/diffstream.py),
short: every format and op, every rounding mode, bits against the reference's
CPU library and the oracle, each element's flags against the oracle's.  And
the generator itself: its GPU build makes the bits of its CPU build."""

import numpy as np
import pytest

import diffstream as ds

pytestmark = pytest.mark.gpu

SIZES = {0: 1 << 14, 1: 1 << 14, 2: 1 << 15, 3: 1 << 16}


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
@pytest.mark.parametrize("op", ["fma", "add", "sub", "mul", "div", "sqrt"])
def test_generator_same_bits_on_gpu_and_cpu(emu, device, fmt, op):
    P = ds.load_probe()
    n, base, seed = 14 * 300, 10**9 + 7, 99
    ga, gb, gc, gf = ds.generate(P, fmt, ds.OPS[op], seed, base, n)
    es = 4 << fmt
    a, b, c = (np.zeros((n, es), np.uint8) for _ in range(3))
    fam = np.zeros(n, np.uint8)
    assert emu.emu_hardcases(fmt, ds.OPS[op], seed, base, n, a.ctypes.data, b.ctypes.data,
                             c.ctypes.data, fam.ctypes.data) == 0
    assert (ga == a).all()
    if gb is not None:
        assert (gb == b).all()
    if gc is not None:
        assert (gc == c).all()
    assert (gf == fam).all()


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
@pytest.mark.parametrize("op", ["fma", "add", "sub", "mul", "div", "sqrt"])
def test_stream_short(device, fmt, op):
    n = SIZES[fmt] if op not in ("div", "sqrt") else SIZES[fmt] // 4
    S = ds.stream(fmt, ds.OPS[op], n, n, seed=2024 + fmt, all_modes=True, dev=device,
                  log=lambda s: print(s, flush=True))
    assert S["reference"], "oracle/_ref/libcft_ref.so is needed for the stream"
    assert S["differ_ref"] == 0 and S["differ_oracle"] == 0, S["first_bad"]
    assert S["oracle_vs_ref"] == 0
    assert S["flag_or_bad"] == 0 and S["flag_groups_bad"] == 0
    assert S["regroup_bits_differ"] == 0
    assert S["clean"]
    assert min(S["fam_cases"]) > 0


@pytest.mark.parametrize("fmt", [1, 2, 3])
@pytest.mark.parametrize("kind", sorted(ds.SEQ_KINDS))
def test_sequencer_stream_short(device, fmt, kind):
    """The same hard cases through one-operation programs in the batched
    runners (per-lane and wide, every mode each), the fused ABS/NEG forms
    included: each lane's result and its own flags against the oracle."""
    chunk = 1 << 12
    S = ds.seq_stream(fmt, kind, 10 * chunk, chunk, seed=77 + fmt, dev=device,
                      log=lambda s: print(s, flush=True))
    assert S["differ"] == 0 and S["flags_differ"] == 0, S["first_bad"]
    assert S["flag_or_bad"] == 0 and S["status_bad"] == 0
    assert sorted(S["by_width"]) == [1, 2]
    assert min(S["per_mode_cases"]) > 0
