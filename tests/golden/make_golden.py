"""Write tests/golden/elem_fp128.npz and elem_fp256.npz: hard-case elementwise
vectors with the reference library's own answers.

Run in the build container, where /root/reference exists and
`make -C oracle ref` has compiled the reference's C library into
oracle/_ref/libcft_ref.so:

    python tests/golden/make_golden.py

Inputs are drawn from fourteen case families (the shape of the reference's
differential streams, host/tools/fx256_stream.h:11-41: random, near 1,
all-ones significands, exact cancellation, constructed ties, subnormals,
overflow and underflow edges, alignment distances at 0, +-1, +-2, p, 2p, 2p+4,
specials).  Expected outputs and per-element flags come from cft_run /
cft_div / cft_sqrt of the reference library, one element per call.  The
fixtures are data: inputs, outputs and flags only.
"""

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "libcft_ref.so")

FMT = {2: (15, 112, 128), 3: (19, 236, 256)}
OPS = {"fma": 0, "add": 1, "sub": 2, "mul": 3, "div": 256, "sqrt": 257}
N = 1024


def enc(fmt, sign, ef, frac):
    ew, mw, w = FMT[fmt]
    return (sign << (w - 1)) | (ef << mw) | (frac & ((1 << mw) - 1))


def gen_inputs(fmt, rng):
    ew, mw, w = FMT[fmt]
    p = mw + 1
    bias = (1 << (ew - 1)) - 1
    emaxf = (1 << ew) - 2
    rb = lambda bits: int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits)
    sg = lambda: int(rng.integers(0, 2))

    def normal(E, frac=None):
        return enc(fmt, sg(), E + bias, rb(mw) if frac is None else frac)

    A, B, C = [], [], []
    fam = 0
    while len(A) < N:
        k = fam % 14
        fam += 1
        if k == 0:      # random over the whole encoding space
            a, b, c = rb(w), rb(w), rb(w)
        elif k == 1:    # near 1.0
            a = normal(int(rng.integers(-2, 3))); b = normal(int(rng.integers(-2, 3)))
            c = normal(int(rng.integers(-4, 5)))
        elif k == 2:    # all-ones significands: rounding carries
            ones = (1 << mw) - 1
            a = normal(int(rng.integers(-3, 3)), ones ^ int(rng.integers(0, 4)))
            b = normal(int(rng.integers(-3, 3)), ones ^ int(rng.integers(0, 4)))
            c = normal(int(rng.integers(-3, 3)), ones)
        elif k == 3:    # exact or near cancellation: c = -(a*b) truncated
            ea, eb = int(rng.integers(-50, 50)), int(rng.integers(-50, 50))
            fa, fb = rb(mw), rb(mw)
            a, b = enc(fmt, 0, ea + bias, fa), enc(fmt, 0, eb + bias, fb)
            P = ((1 << mw) | fa) * ((1 << mw) | fb)
            L = P.bit_length()
            top = P >> (L - p)
            top += int(rng.integers(-2, 3))
            E = ea + eb + (L - 1) - 2 * mw
            c = enc(fmt, 1, E + bias, top)
        elif k == 4:    # ties: product exactly representable + half ulp
            fa = rb(mw // 2) << (mw - mw // 2)
            a = normal(int(rng.integers(-5, 5)), fa)
            b = normal(int(rng.integers(-5, 5)), 0)
            c = normal(int(rng.integers(-5, 5)) - p, rb(mw))
        elif k == 5:    # subnormals
            a = enc(fmt, sg(), 0, rb(mw) >> int(rng.integers(0, mw)))
            b = normal(int(rng.integers(-5, 5)))
            c = enc(fmt, sg(), 0, rb(mw) >> int(rng.integers(0, mw)))
        elif k == 6:    # overflow edge
            a = normal(bias - int(rng.integers(0, 3)))
            b = normal(int(rng.integers(-1, 3)))
            c = normal(bias - int(rng.integers(0, 5)))
        elif k == 7:    # underflow / tininess edge
            a = normal(1 - bias + int(rng.integers(0, 4)))
            b = normal(-int(rng.integers(0, 3)), ((1 << mw) - 1) if rng.random() < 0.5 else None)
            c = enc(fmt, sg(), int(rng.integers(0, 2)), rb(mw))
        elif k in (8, 9, 10, 11):   # alignment distances
            d = [0, 1, -1, 2, -2, p - 1, p, p + 1, -p, 2 * p - 1, 2 * p, 2 * p + 1,
                 2 * p + 3, 2 * p + 4, 2 * p + 5, -(p + 2), -(p + 3), -(p + 4)][
                int(rng.integers(0, 18))]
            ea, eb = int(rng.integers(-40, 40)), int(rng.integers(-40, 40))
            a, b = normal(ea), normal(eb)
            ec = ea + eb - d
            c = normal(max(min(ec, bias), 1 - bias))
        elif k == 12:   # specials against anything
            sp = [enc(fmt, 0, 0, 0), enc(fmt, 1, 0, 0), enc(fmt, 0, (1 << ew) - 1, 0),
                  enc(fmt, 1, (1 << ew) - 1, 0), enc(fmt, 0, (1 << ew) - 1, 1 << (mw - 1)),
                  enc(fmt, 0, (1 << ew) - 1, 1)]
            pick = lambda: sp[int(rng.integers(0, 6))] if rng.random() < 0.6 else rb(w)
            a, b, c = pick(), pick(), pick()
        else:           # wide exponents
            a = enc(fmt, sg(), int(rng.integers(1, emaxf + 1)), rb(mw))
            b = enc(fmt, sg(), int(rng.integers(1, emaxf + 1)), rb(mw))
            c = enc(fmt, sg(), int(rng.integers(1, emaxf + 1)), rb(mw))
        A.append(a); B.append(b); C.append(c)
    esz = w // 8
    pack = lambda v: np.frombuffer(b"".join(x.to_bytes(esz, "little") for x in v),
                                   dtype=np.uint8).copy()
    return pack(A[:N]), pack(B[:N]), pack(C[:N])


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: run `make -C oracle ref` first")
    ref = ctypes.CDLL(REF)
    V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)
    ref.cft_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(V)]
    ref.cft_run.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                            ctypes.c_size_t, U32P, U32P]
    ref.cft_div.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, V, ctypes.c_size_t,
                            U32P, U32P]
    ref.cft_sqrt.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, ctypes.c_size_t, U32P,
                             U32P]
    dev = V()
    assert ref.cft_open(None, 0, ctypes.byref(dev)) == 0
    for fmt, name in ((2, "fp128"), (3, "fp256")):
        rng = np.random.default_rng(20261203 + fmt)
        a, b, c = gen_inputs(fmt, rng)
        esz = FMT[fmt][2] // 8
        out = {"a": a, "b": b, "c": c}
        for opn, op in OPS.items():
            for rnd in range(5):
                d = np.zeros(N * esz, dtype=np.uint8)
                fl = np.zeros(N, dtype=np.uint8)
                f = ctypes.c_uint32()
                bus = ctypes.c_uint32()
                for i in range(N):
                    s = slice(i * esz, (i + 1) * esz)
                    pa = a[s].ctypes.data
                    pb = b[s].ctypes.data
                    pc = c[s].ctypes.data
                    pd = d[s].ctypes.data
                    if op == 256:
                        st = ref.cft_div(dev, fmt, rnd, pa, pb, pd, 1, ctypes.byref(f),
                                         ctypes.byref(bus))
                    elif op == 257:
                        st = ref.cft_sqrt(dev, fmt, rnd, pa, pd, 1, ctypes.byref(f),
                                          ctypes.byref(bus))
                    else:
                        st = ref.cft_run(dev, op, fmt, rnd, pa, pb, pc, pd, 1,
                                         ctypes.byref(f), ctypes.byref(bus))
                    assert st == 0, (name, opn, rnd, i, st)
                    fl[i] = f.value
                out[f"{opn}_r{rnd}"] = d
                out[f"{opn}_r{rnd}_flags"] = fl
        path = os.path.join(HERE, f"elem_{name}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
