"""The sequencer on the CPU: the program emitter and banks against the
reference's certificates, the oracle's executor against the certificates'
boundary states and against the reference library, and the loader's verdicts
(cftx_program_check, the same checks cft_program_load makes) against the
reference loader's."""

import ctypes
import hashlib
import json
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

from cftamd import _lib, check_image
from cftamd import programs as P
from seqlib import ESZ, random_case, ref_device, ref_load, run_oracle, run_ref, diff, \
    rand_values

HERE = os.path.dirname(os.path.abspath(__file__))
CERTS = os.path.join(HERE, "golden", "certs")
REFROOT = "/root/reference"
FMT = {"fp64": 1, "fp256": 3}
STATE_TAG = b"cft-certificate 1 state\x00"


def certs():
    with open(os.path.join(CERTS, "certs.json")) as fh:
        return json.load(fh)


def cert_cases():
    out = []
    for name, c in sorted(certs().items()):
        for r in c["runs"]:
            out.append((name, r["index"]))
    return out


def state(name, run, k):
    with open(os.path.join(CERTS, name, f"run-{run}-boundary-{k}.bin"), "rb") as fh:
        return np.frombuffer(fh.read(), dtype=np.uint8).copy()


def cert_bank(name, system, fmt, run):
    if run == 0:
        return np.frombuffer(P.bank_bytes(system, fmt), dtype=np.uint8).copy()
    with open(os.path.join(CERTS, name, "run-1.bank"), "rb") as fh:
        return np.frombuffer(fh.read(), dtype=np.uint8).copy()


# ---- the emitter and the banks ------------------------------------------------

@pytest.mark.parametrize("name", sorted(certs()))
def test_emitter_matches_certificate_images(name):
    c = certs()[name]
    img = P.GENERATORS[c["system"]](FMT[c["format"]])
    for r in c["runs"]:
        assert hashlib.sha256(img).hexdigest() == r["program_image"]
    ref = os.path.join(REFROOT, "certificates", "images", name + ".cftp")
    if os.path.exists(ref):
        with open(ref, "rb") as fh:
            assert fh.read() == img


@pytest.mark.parametrize("name", sorted(certs()))
def test_banks_match_certificate_digests(name):
    c = certs()[name]
    fmt = FMT[c["format"]]
    img = P.GENERATORS[c["system"]](fmt)
    for r in c["runs"]:
        bank = cert_bank(name, c["system"], fmt, r["index"]).tobytes()
        assert hashlib.sha256(img + bank).hexdigest() == r["program_digest"]
    # the half-step bank is the classic bank at h = 0.005: every slot the
    # certificate halves (H, H2 and H6 or MH) scales exactly by 1/2
    assert P.bank_bytes(c["system"], fmt, "0.005") == cert_bank(name, c["system"], fmt,
                                                                 1).tobytes()


def test_round_rational_against_python_floats():
    rng = np.random.default_rng(7)
    texts = ["0.01", "0.001", "0.1", "1e-3", "2.5", "8", "28", "10", "3.14159",
             "6.02214076e23", "1e-300", "0.3"]
    for _ in range(300):
        texts.append(f"{rng.integers(1, 10**9)}e{rng.integers(-40, 40)}")
    for t in texts:
        assert P.dec(1, t) == struct.unpack("<Q", struct.pack("<d", float(t)))[0], t
    # 8/3 and 1/6 in binary64
    assert P.round_rational(1, Fraction(8, 3)) == struct.unpack("<Q", struct.pack("<d", 8 / 3))[0]


@pytest.mark.parametrize("fmt", [2, 3])
def test_round_rational_is_nearest_even(fmt):
    """Wide formats: the rounded value lies within half an ulp, and exact
    halfway cases go to the even neighbour."""
    width, p, ew = P.FORMATS[fmt]
    rng = np.random.default_rng(fmt)
    for _ in range(200):
        q = Fraction(int(rng.integers(1, 2**62)), int(rng.integers(1, 2**62)))
        v = P.round_rational(fmt, q)
        x = P.to_rational(fmt, v)
        e = (v >> (p - 1)) & ((1 << ew) - 1)
        ulp = Fraction(2) ** (e - ((1 << (ew - 1)) - 1) - (p - 1))
        assert abs(x - q) <= ulp / 2
    one = P.round_rational(fmt, Fraction(1))
    ulp1 = Fraction(1, 2 ** (p - 1))
    assert P.round_rational(fmt, 1 + ulp1 / 2) == one                 # tie to even (down)
    assert P.round_rational(fmt, 1 + 3 * ulp1 / 2) == one + 2         # tie to even (up)


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_round_sqrt_brackets_the_root(fmt):
    """RN(sqrt(q)) lies between the midpoints to its neighbours, squared (the
    lower neighbour half as far at a power of two); perfect squares are
    exact; binary64 agrees with the platform's correctly rounded sqrt."""
    import math
    width, p, ew = P.FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    rng = np.random.default_rng(40 + fmt)
    for _ in range(300):
        q = Fraction(int(rng.integers(1, 2**62)), int(rng.integers(1, 2**62))) * \
            Fraction(2) ** int(rng.integers(-30, 31))
        v = P.round_sqrt(fmt, q)
        r = P.to_rational(fmt, v)
        e = (v >> (p - 1)) & ((1 << ew) - 1)
        ulp = Fraction(2) ** (e - bias - (p - 1))
        below = ulp / 2 if (v & ((1 << (p - 1)) - 1)) == 0 else ulp
        assert (r - below / 2) ** 2 <= q <= (r + ulp / 2) ** 2, q
        if fmt == 1:
            assert v == struct.unpack("<Q", struct.pack("<d", math.sqrt(float(q))))[0] or \
                float(q) != q
    for k in (k for k in (1, 3, 12345, 2**40 + 1) if k.bit_length() <= p):
        assert P.to_rational(fmt, P.round_sqrt(fmt, Fraction(k * k, 4 ** 7))) == \
            Fraction(k, 2 ** 7)


def test_henon_heiles_orbits_at_the_energy():
    """Orbit states carry E = 1/8 to binary256 precision: px is the
    correctly rounded root, so H(x, y, px, py) - 1/8 is within a few ulp of
    px^2 / 2 (about 2^-237 relative)."""
    from cftamd import odes
    _img, _bank, st, _steps = odes.henonheiles(8, seed=3, fmt=3)
    for row in st.reshape(8, 4, 32):
        x, y, px, py = (P.to_rational(3, int.from_bytes(v.tobytes(), "little")) for v in row)
        H = (px * px + py * py) / 2 + (x * x + y * y) / 2 + x * x * y - y ** 3 / 3
        assert abs(H - Fraction(1, 8)) <= Fraction(1, 2 ** 230)


# ---- the oracle's executor ---------------------------------------------------------

def cert_run(oracle_or_none, name, run, runner):
    """Every segment of one certificate run through `runner`; returns the
    first disagreement or None."""
    c = certs()[name]
    fmt = FMT[c["format"]]
    sysname = c["system"]
    r = [x for x in c["runs"] if x["index"] == run][0]
    img = P.GENERATORS[sysname](fmt)
    bank = cert_bank(name, sysname, fmt, run)
    lanes = r["lanes"]
    esz = ESZ[fmt]
    for k, seg in enumerate(r["segments"]):
        start = state(name, run, k)
        assert hashlib.sha256(STATE_TAG + start.tobytes()).hexdigest() == seg["start"]
        slots = start.nbytes // (lanes * esz)
        case = {"image": img, "bank": bank, "a": np.zeros(lanes * esz, dtype=np.uint8),
                "b": None, "c": None, "n": lanes, "fmt": fmt, "scratch_in": start,
                "n_out": slots, "max_dep": 0, "mask": None}
        out = runner(case)
        end = out["scratch_out"]
        if hashlib.sha256(STATE_TAG + end.tobytes()).hexdigest() != seg["end"]:
            return f"segment {k}: end state differs"
        if out["flags"] != seg["flags"] or out["status"] != seg["status"]:
            return f"segment {k}: flags {out['flags']} status {out['status']}"
        if not np.array_equal(end, state(name, run, k + 1)):
            return f"segment {k}: end state differs from the boundary file"
    return None


@pytest.mark.parametrize("name,run", cert_cases())
def test_oracle_reproduces_certificates(oracle, name, run):
    assert cert_run(oracle, name, run, lambda case: run_oracle(oracle, case)) is None


def test_oracle_matches_reference_on_random_programs(oracle, ref):
    rng = np.random.default_rng(2024)
    for it in range(240):
        fmt = int(rng.integers(0, 4))
        case = random_case(rng, fmt, int(rng.integers(1, 80)), budget=int(rng.integers(10, 90)))
        x = run_oracle(oracle, case)
        y = run_ref(ref, case)
        assert diff(x, y) == [], (it, fmt)


@pytest.mark.parametrize("fmt", [1, 3])
def test_oracle_matches_reference_on_odes(oracle, ref, fmt):
    rng = np.random.default_rng(fmt)
    for sysname, kw, slots in (("lorenz63-rk4", {}, 3), ("henonheiles-lf", {}, 4),
                               ("lorenz96-rk4", {"k": 9}, 9), ("lorenz96-rk4", {"k": 16}, 16)):
        img = P.with_trips(P.GENERATORS[sysname](fmt, **kw), 7)
        lanes = 40
        bank = np.frombuffer(P.bank_bytes(sysname, fmt), dtype=np.uint8).copy()
        sin = np.concatenate([np.frombuffer(struct.pack("<d", 1 + rng.uniform(-1e-3, 1e-3)),
                                            dtype=np.uint8) for _ in range(lanes * slots)])
        if fmt == 3:      # widen the binary64 draws to binary256 exactly
            sin = widen64(sin)
        case = {"image": img, "bank": bank, "a": np.zeros(lanes * ESZ[fmt], dtype=np.uint8),
                "b": None, "c": None, "n": lanes, "fmt": fmt, "scratch_in": sin,
                "n_out": slots, "max_dep": 0, "mask": None}
        assert diff(run_oracle(oracle, case), run_ref(ref, case)) == [], sysname


def widen64(b):
    """binary64 encodings -> the same values in binary256 (exact)."""
    out = []
    for (v,) in struct.iter_unpack("<Q", b.tobytes()):
        s, e, f = v >> 63, (v >> 52) & 0x7FF, v & ((1 << 52) - 1)
        assert 0 < e < 0x7FF
        E = e - 1023 + 262143
        out.append(((s << 255) | (E << 236) | (f << 184)).to_bytes(32, "little"))
    return np.frombuffer(b"".join(out), dtype=np.uint8).copy()


# ---- the loader ------------------------------------------------------------------

def mutate(rng, img):
    b = bytearray(img)
    for _ in range(int(rng.integers(1, 4))):
        r = rng.random()
        if r < 0.15 and len(b) > 32:
            b = b[:int(rng.integers(0, len(b)))]
        elif r < 0.3 and len(b) >= 32:
            i = int(rng.integers(0, 32))
            b[i] ^= 1 << int(rng.integers(0, 8))
        elif len(b):
            i = int(rng.integers(0, len(b)))
            b[i] ^= 1 << int(rng.integers(0, 8))
    return bytes(b)


def has_aug(img):
    if len(img) < 32:
        return False
    n_insns, n_consts, _md, fmt, flags = struct.unpack_from("<5I", img, 8)
    if fmt > 3:
        return False
    off = 32 + (0 if flags & 1 else n_consts * ESZ[fmt])
    for i in range(n_insns):
        if off + 8 * i + 8 > len(img):
            return False
        w, = struct.unpack_from("<Q", img, off + 8 * i)
        if (w >> 31) & 1 and (w & 0xFF) in (10, 11):
            return True
    return False


def hand_cases(depth):
    """Images that each break one loader rule."""
    L = P.lorenz63(3)
    good = [P.alu(0, 1, 2, 3, 4), P.halt()]
    bad = {
        "short": L[:20],
        "magic": b"XFTP" + L[4:],
        "version": L[:4] + struct.pack("<I", 2) + L[8:],
        "unknown flag": L[:24] + struct.pack("<I", 3 | 8) + L[28:],
        "scratch_io without flag": P.image(3, good, scratch_io=0) [:28] + struct.pack("<I", 5)
                                   + P.image(3, good)[32:],
        "precision": L[:20] + struct.pack("<I", 4) + L[24:],
        "length": L + b"\0" * 8,
        "deposit cap": P.image(3, good, max_deposits=(1 << 20) + 1),
        "513 constants": P.image(1, good, n_consts=513, bank_ext=True),
        "scratch in past depth": P.image(3, good, scratch_in=depth + 1, scratch_out=1),
        "kx selects nothing": P.image(3, [P.alu(0, 1, 2, 3, 4, kx=True), P.halt()]),
        "imm31": P.image(3, [P.alu(0, 1, 2, 3, 4) | (1 << 63), P.halt()]),
        "const index": P.image(3, [P.alu(0, 1, P.K(3), 2, 3), P.halt()], consts=[0, 0]),
        "rnd 5": P.image(3, [P.alu(0, 1, 2, 3, 4, rnd=5), P.halt()]),
        "ctrl 15": P.image(3, [P.ctrl(15), P.halt()]),
        "ctrl reads rc": P.image(3, [P.ctrl(0) | (3 << 20)]),
        "repeat 0": P.image(3, [P.repeat(0), P.endrep()]),
        "five loops": P.image(3, [P.repeat(2)] * 5 + [P.endrep()] * 5),
        "endrep alone": P.image(3, [P.endrep()]),
        "actall in loop": P.image(3, [P.repeat(2), P.actall(), P.endrep()]),
        "halt in loop": P.image(3, [P.repeat(2), P.halt(), P.endrep()]),
        "open loop": P.image(3, [P.repeat(2)]),
        "open quiet": P.image(3, [P.quiet()]),
        "crossed brackets": P.image(3, [P.repeat(2), P.quiet(), P.endrep(), P.endquiet()]),
        "halt in quiet": P.image(3, [P.quiet(), P.halt(), P.endquiet()]),
        "worst case": P.image(3, [P.repeat(1 << 20), P.repeat(1 << 20), P.alu(0, 1, 2, 3, 4),
                                  P.endrep(), P.endrep()]),
        "static slot": P.image(3, [P.stl(1, depth)]),
        "augadd": P.image(3, [P.ctrl(10, rd=1, ra=2, rb=3)]),
    }
    return bad


@pytest.mark.parametrize("depth", [256, 2048])
def test_loader_matches_reference_on_rule_breakers(ref, depth):
    dev = ref_device(ref, depth)
    try:
        for what, img in hand_cases(depth).items():
            mine, _ = check_image(img, depth)
            theirs, h = ref_load(ref, dev, img)
            if theirs == 0:
                ref.cft_program_free(h)
            if what == "augadd":
                assert theirs == 0 and mine == _lib.ERR_UNSUPPORTED, what
            else:
                assert theirs != 0, what
                assert mine == theirs, (what, mine, theirs)
    finally:
        ref.cft_close(dev)


def test_loader_matches_reference_on_mutations(ref):
    rng = np.random.default_rng(99)
    dev = ref_device(ref, 256)
    seen = set()
    try:
        for it in range(1500):
            fmt = int(rng.integers(0, 4))
            img = random_case(rng, fmt, 1, budget=int(rng.integers(3, 30)))["image"]
            if rng.random() < 0.85:
                img = mutate(rng, img)
            mine, info = check_image(img, 256)
            theirs, h = ref_load(ref, dev, img)
            if theirs == 0:
                ref.cft_program_free(h)
            if theirs == 0 and has_aug(img):
                assert mine == _lib.ERR_UNSUPPORTED
                continue
            assert mine == theirs, (it, mine, theirs, _lib.lib().cft_last_error())
            seen.add(theirs)
    finally:
        ref.cft_close(dev)
    assert {0, _lib.ERR_INVALID_ARGUMENT, _lib.ERR_ARTIFACT} <= seen


def test_loader_info_matches_reference(ref):
    ref.cft_program_get_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.CftProgramInfo)]
    for depth in (256, 2048):
        dev = ref_device(ref, depth)
        for img in (P.lorenz63(3), P.lorenz96(1), P.henonheiles(3), P.lorenz96(3, k=48),
                    P.image(2, [P.stx(1, 2), P.halt()], max_deposits=3)):
            st, mine = check_image(img, depth)
            theirs = _lib.CftProgramInfo()
            theirs.struct_size = ctypes.sizeof(theirs)
            rs, h = ref_load(ref, dev, img)
            assert st == rs == 0
            assert ref.cft_program_get_info(h, ctypes.byref(theirs)) == 0
            ref.cft_program_free(h)
            for f, _t in _lib.CftProgramInfo._fields_:
                assert getattr(mine, f) == getattr(theirs, f), f
        ref.cft_close(dev)
