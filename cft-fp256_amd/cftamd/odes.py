"""Seeded ODE ensembles for the batched runner: Lorenz-63 and Lorenz-96 RK4
and Henon-Heiles leapfrog, as (program image, bank, initial states, steps).

Initial perturbations are dyadic rationals (an integer times 2^-40 for
Lorenz-63, 2^-36 for Lorenz-96), so 1 + delta and F + delta are exact binary64
values and identical in every wider format.  Element encodings are built
from binary64 values by exact widening.
"""

import numpy as np

from . import programs as P

ESZ = {1: 8, 2: 16, 3: 32}


def widen(x64, fmt):
    """binary64 values (finite, normal or zero) -> dense encodings of fmt."""
    x = np.ascontiguousarray(x64, dtype=np.float64)
    if fmt == 1:
        return x.view(np.uint8).copy()
    width, p, ew = P.FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    u = x.view(np.uint64)
    sign = (u >> np.uint64(63)).astype(object)
    e = ((u >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int64)
    f = (u & np.uint64((1 << 52) - 1)).astype(object)
    if np.any((e == 0) & (f != 0)) or np.any(e == 0x7FF):
        raise ValueError("only normal binary64 values and zeros widen here")
    esz = width // 8
    out = bytearray(x.size * esz)
    shift = p - 53
    for i in range(x.size):
        if e[i] == 0:
            v = int(sign[i]) << (width - 1)
        else:
            v = (int(sign[i]) << (width - 1)) | ((int(e[i]) - 1023 + bias) << (p - 1)) | \
                (int(f[i]) << shift)
        out[i * esz:(i + 1) * esz] = v.to_bytes(esz, "little")
    return np.frombuffer(bytes(out), dtype=np.uint8).copy()


def _dyadic(rng, n, bound, bits):
    k = int(round(bound * 2.0 ** bits))
    return rng.integers(-k, k + 1, size=n).astype(np.float64) * 2.0 ** -bits


def lorenz63(n, seed=0, fmt=3, steps_lo=1000, steps_hi=100000, h="0.001", seg=100):
    """n trajectories from (1,1,1) + delta, delta dyadic in [-1e-3, 1e-3];
    steps log-uniform in [steps_lo, steps_hi]; the classic bank with h."""
    rng = np.random.default_rng(seed)
    x = 1.0 + _dyadic(rng, 3 * n, 1e-3, 40)
    steps = np.exp(rng.uniform(np.log(steps_lo), np.log(steps_hi), size=n))
    steps = np.clip(np.rint(steps), steps_lo, steps_hi).astype(np.uint32)
    img = P.with_trips(P.lorenz63(fmt), seg)
    bank = np.frombuffer(P.bank_bytes("lorenz63-rk4", fmt, h), dtype=np.uint8).copy()
    return img, bank, widen(x, fmt), steps


def lorenz96(n, seed=0, fmt=3, steps_lo=1000, steps_hi=20000, ks=None, forcing=8.0,
             h="0.01", seg=20):
    """n instances with ring size K uniform in ks (default 8, 16, ..., 256),
    x_i = F + delta, delta dyadic in [-0.01, 0.01]; steps uniform in
    [steps_lo, steps_hi].  Returns a list of groups, one per K:
    (k, image, bank, lane ids, states, steps)."""
    rng = np.random.default_rng(seed)
    ks = list(range(8, 257, 8)) if ks is None else list(ks)
    kk = rng.choice(ks, size=n)
    steps = rng.integers(steps_lo, steps_hi + 1, size=n).astype(np.uint32)
    bank = np.frombuffer(P.bank_bytes("lorenz96-rk4", fmt, h), dtype=np.uint8).copy()
    groups = []
    for k in sorted(set(kk.tolist())):
        ids = np.flatnonzero(kk == k)
        x = forcing + _dyadic(rng, k * ids.size, 0.01, 36)
        img = P.with_trips(P.lorenz96(fmt, k=k), seg)
        groups.append((k, img, bank, ids, widen(x, fmt), steps[ids]))
    return groups


def henonheiles(n, seed=0, fmt=3, steps_lo=10000, steps_hi=100000, h="0.01", seg=100,
                energy="1/8"):
    """n orbits at energy E: x, y, py drawn uniform in [-0.2, 0.2] (binary64
    values, exact in every wider format), px = RN(sqrt(2E - py^2 - x^2 - y^2
    - 2 x^2 y + 2 y^3 / 3)) in the format itself, the argument evaluated
    exactly as a rational and its square root correctly rounded, so the
    orbit's energy is E to the format's own precision."""
    from fractions import Fraction
    rng = np.random.default_rng(seed)
    esz = ESZ[fmt]
    E = Fraction(energy)
    rows = []
    while len(rows) < n:
        x, y, py = rng.uniform(-0.2, 0.2, size=3)
        fx, fy, fpy = Fraction(float(x)), Fraction(float(y)), Fraction(float(py))
        rest = 2 * E - fpy * fpy - fx * fx - fy * fy - 2 * fx * fx * fy + 2 * fy ** 3 / 3
        if rest <= 0:
            continue
        enc = widen(np.array([x, y, py]), fmt).reshape(3, esz)
        px = np.frombuffer(P.round_sqrt(fmt, rest).to_bytes(esz, "little"), dtype=np.uint8)
        rows.append(np.concatenate([enc[0], enc[1], px, enc[2]]))
    steps = rng.integers(steps_lo, steps_hi + 1, size=n).astype(np.uint32)
    img = P.with_trips(P.henonheiles(fmt), seg)
    bank = np.frombuffer(P.bank_bytes("henonheiles-lf", fmt, h), dtype=np.uint8).copy()
    return img, bank, np.concatenate(rows).copy(), steps
