"""The wide scheduler on the CPU: cftx_program_wide's step records, executed
by a plain simulation of the wide kernel's semantics (one instance, its
lanes' records in any order, operands' sign modifiers applied on load, the
oracle computing each ALU operation), give
the same scratch-out block and flags as the oracle's ordinary run of the
program - for the ODE programs at every width and for random pure
programs."""

import ctypes
import struct

import numpy as np
import pytest

from cftamd import _lib, odes
from cftamd import programs as P
from seqlib import ESZ, rand_values, run_oracle

WK_ALU, WK_LOAD, WK_STORE, WK_MOVE, WK_NOP, WK_CTRL = range(6)


def wide(image, width, depth=256, max_vregs=0):
    L = _lib.lib()
    im = np.frombuffer(bytes(image), dtype=np.uint8).copy()
    info = (ctypes.c_uint32 * 6)()
    st = L.cftx_program_wide(im.ctypes.data, im.nbytes, depth, width, max_vregs, None, 0, info)
    if st != 0:
        return st, None, None
    recs = np.zeros(info[0] * width * 4, dtype=np.uint32)
    st = L.cftx_program_wide(im.ctypes.data, im.nbytes, depth, width, max_vregs,
                             recs.ctypes.data, info[0] * width, info)
    return st, recs.reshape(info[0], width, 4), list(info)


def simulate(oracle, recs, info, fmt, bank, sin, n_out, trips, a=None, phys_a=-1):
    """One instance through the steps; returns (scratch-out bytes, flags)."""
    esz = ESZ[fmt]
    nsteps, nvregs, nregs, top = info[0], info[1], info[2], info[3]
    zero = bytes(esz)
    regs = [zero] * nvregs
    if a is not None and phys_a >= 0:
        regs[phys_a] = a
    scratch = {}
    n_in = len(sin) // esz
    for s in range(n_in):
        scratch[s] = sin[s * esz:(s + 1) * esz]
    konst = [bank[i * esz:(i + 1) * esz] for i in range(len(bank) // esz)] if bank else []
    out = ctypes.create_string_buffer(esz)
    f = ctypes.c_uint32()
    flags = 0
    stack = []
    pc = 0
    while pc < nsteps:
        r0 = recs[pc][0]
        kind0 = (int(r0[0]) >> 8) & 7
        if kind0 == WK_CTRL:
            op = int(r0[0]) & 0xFF
            if op == 0:
                break
            if op == 1:                                  # REPEAT
                t = trips if pc == top else int(r0[3])
                if t == 0:
                    pc = int(r0[1]) + 1
                    continue
                stack.append([pc + 1, t])
                pc += 1
            else:                                        # ENDREP
                stack[-1][1] -= 1
                if stack[-1][1] > 0:
                    pc = stack[-1][0]
                else:
                    stack.pop()
                    pc += 1
            continue
        writes = []
        for rec in recs[pc]:
            x, y, z = int(rec[0]), int(rec[1]), int(rec[2])
            kind = (x >> 8) & 7
            if kind == WK_NOP:
                continue
            if kind == WK_ALU:
                def src(v, k, m):
                    e = bytearray(konst[v] if k else regs[v])
                    if m & 1:                            # a fused ABS
                        e[-1] &= 0x7F
                    if m & 2:                            # a fused NEG
                        e[-1] ^= 0x80
                    return bytes(e)
                sa = src(y & 0xFFFF, (x >> 14) & 1, (x >> 17) & 3)
                sb = src(y >> 16, (x >> 15) & 1, (x >> 19) & 3)
                sc = src(z & 0xFFFF, (x >> 16) & 1, (x >> 21) & 3)
                oracle.oc_compute(fmt, x & 0xFF, (x >> 11) & 7, sa, sb, sc, out,
                                  ctypes.byref(f))
                flags |= f.value
                writes.append(("r", z >> 16, out.raw))
            elif kind == WK_LOAD:
                writes.append(("r", z >> 16, scratch.get(y & 0xFFFF, zero)))
            elif kind == WK_STORE:
                writes.append(("s", z >> 16, regs[y & 0xFFFF]))
            else:
                writes.append(("r", z >> 16, regs[y & 0xFFFF]))
        for where, i, v in writes:
            if where == "r":
                regs[i] = v
            else:
                scratch[i] = v
        pc += 1
    return b"".join(scratch.get(s, zero) for s in range(n_out)), flags


def oracle_one(oracle, img, fmt, bank, sin, n_out, trips):
    case = {"image": P.with_trips(img, trips), "bank": None if bank is None else
            np.frombuffer(bank, dtype=np.uint8).copy(), "a": np.zeros(ESZ[fmt], np.uint8),
            "b": None, "c": None, "n": 1, "fmt": fmt,
            "scratch_in": np.frombuffer(sin, dtype=np.uint8).copy(), "n_out": n_out,
            "max_dep": 0, "mask": None}
    o = run_oracle(oracle, case, threads=1)
    return o["scratch_out"].tobytes(), o["flags"]


@pytest.mark.parametrize("system,fmt,k,widths", [
    ("lorenz63-rk4", 3, None, [1, 2, 4, 8]), ("henonheiles-lf", 3, None, [1, 2, 4]),
    ("lorenz96-rk4", 1, 12, [1, 4, 16, 64]), ("lorenz63-rk4", 1, None, [4])])
def test_ode_schedules_match_the_oracle(oracle, system, fmt, k, widths):
    oracle.oc_compute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                  ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_uint32)]
    kw = {"k": k} if k else {}
    img = P.GENERATORS[system](fmt, **kw)
    slots = {"lorenz63-rk4": 3, "henonheiles-lf": 4}.get(system, k)
    bank = P.bank_bytes(system, fmt)
    rng = np.random.default_rng(1)
    sin = odes.widen(1 + rng.uniform(-1e-3, 1e-3, size=slots), fmt).tobytes()
    want = oracle_one(oracle, img, fmt, bank, sin, slots, 3)
    for w in widths:
        st, recs, info = wide(img, w)
        assert st == 0
        got = simulate(oracle, recs, info, fmt, bank, sin, slots, 3)
        assert got == want, (system, w)


def pure_program(rng, fmt, n_io, nk):
    """A random program of ALU instructions and static scratch accesses in
    nested loops, its state in slots [0, n_io)."""
    ins = []

    def block(depth, budget):
        while budget > 0:
            r = rng.random()
            budget -= 1
            if r < 0.6:
                op = int(rng.choice([0, 0, 1, 2, 3, 3, 4, 5, 6, 7, 8, 11, 16, 19, 20, 21, 26]))
                srcs = [P.K(int(rng.integers(0, nk))) if nk and rng.random() < 0.3
                        else int(rng.integers(0, 12)) for _ in range(3)]
                ins.append(P.alu(op, int(rng.integers(0, 12)), *srcs,
                                 rnd=int(rng.integers(0, 5))))
            elif r < 0.75:
                ins.append(P.ldl(int(rng.integers(0, 12)), int(rng.integers(0, n_io + 4))))
            elif r < 0.9:
                ins.append(P.stl(int(rng.integers(0, 12)), int(rng.integers(0, n_io + 4))))
            elif depth < 3 and budget > 3:
                ins.append(P.repeat(int(rng.integers(1, 4))))
                used = int(rng.integers(2, min(budget, 10) + 1))
                block(depth + 1, used)
                ins.append(P.endrep())
                budget -= used
    ins.append(P.repeat(2))
    block(1, int(rng.integers(5, 60)))
    ins.append(P.endrep())
    block(0, int(rng.integers(0, 8)))
    ins.append(P.halt())
    consts = [int.from_bytes(rand_values(rng, fmt, 1).tobytes(), "little") for _ in range(nk)]
    return P.image(fmt, ins, consts=consts, scratch_in=n_io, scratch_out=n_io), consts


def test_random_pure_programs_match_the_oracle(oracle):
    oracle.oc_compute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                  ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_uint32)]
    rng = np.random.default_rng(77)
    for it in range(60):
        fmt = int(rng.integers(0, 4))
        n_io = int(rng.integers(1, 6))
        img, consts = pure_program(rng, fmt, n_io, int(rng.integers(0, 20)))
        esz = ESZ[fmt]
        bank = b"".join(v.to_bytes(esz, "little") for v in consts)
        sin = rand_values(rng, fmt, n_io).tobytes()
        trips = int(rng.integers(1, 4))
        want = oracle_one(oracle, img, fmt, None, sin, n_io, trips)
        for w in (1, 2, 4, 16):
            st, recs, info = wide(img, w)
            assert st == 0, it
            got = simulate(oracle, recs, info, fmt, bank, sin, n_io, trips)
            assert got == want, (it, fmt, w)


def test_impure_programs_are_refused():
    for img in (P.image(3, [P.deposit(1), P.halt()], max_deposits=1),
                P.image(3, [P.setact(1), P.halt()]),
                P.image(3, [P.stx(1, 2), P.halt()]),
                P.image(3, [P.quiet(), P.endquiet()])):
        st, _r, _i = wide(img, 4)
        assert st == _lib.ERR_UNSUPPORTED
