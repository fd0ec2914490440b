"""Test helpers for the sequencer: random valid programs, operands, and the
three ways to run a program - the oracle restatement (oracle/seq_oracle.c),
the reference's own library where it is built (oracle/_ref), and libcft on
the GPU - each returning the same dictionary of outputs."""

import ctypes
import struct

import numpy as np

from cftamd import programs as P

ESZ = {0: 4, 1: 8, 2: 16, 3: 32}
V = ctypes.c_void_p
U32P = ctypes.POINTER(ctypes.c_uint32)

# opcodes a program may issue: the arithmetic four weighted up, every other
# assigned one, a few reductions and unassigned codes (qNaN + invalid)
ARITH = [0, 1, 2, 3]
MISC = [4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 16, 17, 18, 19, 20, 21, 22, 23, 26, 27, 30]
ODD = [15, 24, 25, 31, 77, 255]


def rand_values(rng, fmt, n):
    """n encodings of format fmt, mixing specials, small integers, values near
    1 and random bit patterns; as a dense little-endian uint8 array."""
    width, p, ew = P.FORMATS[fmt]
    bias = (1 << (ew - 1)) - 1
    esz = width // 8
    out = np.empty(n * esz, dtype=np.uint8)
    top = lambda s, e, f: (s << (width - 1)) | (e << (p - 1)) | f
    for i in range(n):
        r = rng.random()
        s = int(rng.integers(0, 2))
        if r < 0.08:
            v = [0, 1 << (width - 1), top(0, (1 << ew) - 1, 0), top(1, (1 << ew) - 1, 0),
                 top(0, (1 << ew) - 1, 1 << (p - 2)), top(0, (1 << ew) - 1, 1),
                 top(0, 0, 1)][int(rng.integers(0, 7))]
        elif r < 0.2:
            v = int(rng.integers(0, 300))          # a small integer bit pattern
        elif r < 0.7:
            e = bias + int(rng.integers(-8, 9))
            f = int.from_bytes(rng.bytes(esz), "little") & ((1 << (p - 1)) - 1)
            v = top(s, e, f)
        else:
            v = int.from_bytes(rng.bytes(esz), "little")
        out[i * esz:(i + 1) * esz] = np.frombuffer(v.to_bytes(esz, "little"), dtype=np.uint8)
    return out


class Gen:
    """A random program the loader accepts: nested loops and quiet regions,
    every control code but AUGADD/AUGERR, constants plain and under kx."""

    def __init__(self, rng, fmt, n_consts, slots=24, depth=256, deposits=True,
                 scratch=True, flag_control=True):
        self.rng, self.fmt, self.nk = rng, fmt, n_consts
        self.slots, self.depth = slots, depth
        self.deposits, self.scratch, self.fc = deposits, scratch, flag_control
        self.P = []

    def reg(self):
        return int(self.rng.integers(0, 32))

    def src(self):
        if self.nk and self.rng.random() < 0.3:
            return P.K(int(self.rng.integers(0, self.nk)))
        return self.reg()

    def alu(self):
        r = self.rng.random()
        op = (int(self.rng.choice(ARITH)) if r < 0.55 else
              int(self.rng.choice(MISC)) if r < 0.95 else int(self.rng.choice(ODD)))
        rnd = int(self.rng.integers(0, 5))
        a, b, c = self.src(), self.src(), self.src()
        kx = None
        if any(isinstance(s, P.Const) for s in (a, b, c)) and self.rng.random() < 0.2:
            kx = True
        self.P.append(P.alu(op, self.reg(), a, b, c, rnd, kx=kx))

    def index_reg(self):
        """A register made to hold a small integer (an in-range slot mostly)."""
        rb = self.reg()
        if self.nk >= 2 and self.rng.random() < 0.8:
            self.P.append(P.alu(16, rb, self.reg(), P.K(self.nk - 1)))   # iand with a mask
        return rb

    def block(self, loops, quiets, budget):
        while budget > 0:
            r = self.rng.random()
            budget -= 1
            if r < 0.55:
                self.alu()
            elif r < 0.62 and loops < 4 and budget > 4:
                n = int(self.rng.integers(1, 4))
                self.P.append(P.repeat(n))
                used = int(self.rng.integers(2, min(budget, 12) + 1))
                self.block(loops + 1, quiets, used)
                self.P.append(P.endrep())
                budget -= used
            elif r < 0.66 and self.fc and quiets < 4 and budget > 3:
                self.P.append(P.quiet())
                used = int(self.rng.integers(1, min(budget, 8) + 1))
                self.block(loops, quiets + 1, used)
                self.P.append(P.endquiet())
                budget -= used
            elif r < 0.70:
                self.P.append(P.setact(self.reg()))
            elif r < 0.72 and loops == 0:
                self.P.append(P.actall())
            elif r < 0.77 and self.deposits:
                self.P.append(P.deposit(self.reg()))
            elif r < 0.82 and self.scratch:
                slot = int(self.rng.integers(0, self.slots))
                self.P.append(P.stl(self.reg(), slot) if self.rng.random() < 0.5
                              else P.ldl(self.reg(), slot))
            elif r < 0.90 and self.scratch:
                rb = self.index_reg()
                step = int(self.rng.integers(-3, 4)) if self.rng.random() < 0.5 else 0
                self.P.append(P.stx(self.reg(), rb, step) if self.rng.random() < 0.5
                              else P.ldx(self.reg(), rb, step))
            elif r < 0.95 and self.fc:
                self.P.append(P.raise_(self.reg()))
            else:
                self.alu()

    def build(self, budget=60, halt=True):
        self.block(0, 0, budget)
        if halt:
            self.P.append(P.halt())
            if self.rng.random() < 0.3:
                self.alu()                   # unreachable tail
        return self.P


def random_case(rng, fmt, n, budget=60, depth=256):
    """A random program image, its bank and the run's inputs."""
    nk = int(rng.integers(0, 24))
    g = Gen(rng, fmt, nk, depth=depth)
    insns = g.build(budget)
    bank_ext = bool(rng.integers(0, 2))
    strict = bool(rng.integers(0, 2))
    n_in = int(rng.integers(0, 6))
    n_out = int(rng.integers(0, 6))
    max_dep = int(rng.integers(0, 6))
    width = P.FORMATS[fmt][0]
    consts = [int.from_bytes(rand_values(rng, fmt, 1).tobytes(), "little") for _ in range(nk)]
    if nk >= 2:
        consts[-1] = (1 << int(rng.integers(3, 12))) - 1      # the index mask
    esz = width // 8
    kbytes = b"".join(v.to_bytes(esz, "little") for v in consts)
    img = P.image(fmt, insns, consts=consts, n_consts=nk, max_deposits=max_dep,
                  bank_ext=bank_ext, strict=strict, scratch_io=n_in | (n_out << 16))
    bank = np.frombuffer(kbytes, dtype=np.uint8).copy() if bank_ext and nk else None
    a = rand_values(rng, fmt, n)
    b = rand_values(rng, fmt, n) if rng.random() < 0.8 else None
    c = rand_values(rng, fmt, n) if rng.random() < 0.8 else None
    sin = rand_values(rng, fmt, n * n_in) if n_in else None
    mask = None
    if rng.random() < 0.3:
        mask = rng.integers(0, 256, size=(n + 7) // 8, dtype=np.uint8)
    return {"image": img, "bank": bank, "a": a, "b": b, "c": c, "n": n, "fmt": fmt,
            "scratch_in": sin, "n_out": n_out, "max_dep": max_dep, "mask": mask}


def _outputs(case, sentinel=0xA5):
    n, esz = case["n"], ESZ[case["fmt"]]
    return {
        "deposits": np.full(n * case["max_dep"] * esz, sentinel, dtype=np.uint8)
        if case["max_dep"] else None,
        "counts": np.full(n, 0xDEADBEEF, dtype=np.uint32),
        "lane_flags": np.full(n, sentinel, dtype=np.uint8),
        "scratch_out": np.full(n * case["n_out"] * esz, sentinel, dtype=np.uint8)
        if case["n_out"] else None,
    }


def _p(x):
    return None if x is None else x.ctypes.data


def run_oracle(oracle, case, depth=256, threads=8):
    out = _outputs(case)
    fl, st = ctypes.c_uint32(), ctypes.c_uint32()
    img = np.frombuffer(case["image"], dtype=np.uint8)
    rc = oracle.oc_seq_run(img.ctypes.data, img.nbytes, _p(case["bank"]), _p(case["a"]),
                           _p(case["b"]), _p(case["c"]), case["n"], _p(case["scratch_in"]),
                           _p(out["scratch_out"]), _p(out["deposits"]),
                           out["counts"].ctypes.data_as(U32P), _p(case["mask"]),
                           _p(out["lane_flags"]), depth, threads, ctypes.byref(fl),
                           ctypes.byref(st))
    assert rc == 0, "the oracle refused the case"
    out["flags"], out["status"] = fl.value, st.value
    return out


class RunArgs(ctypes.Structure):
    """The reference's cft_run_args (host/include/cft.h:2935-3052)."""
    _fields_ = [("struct_size", ctypes.c_size_t), ("a", V), ("b", V), ("c", V),
                ("n", ctypes.c_size_t), ("bank", V), ("bank_bytes", ctypes.c_size_t),
                ("scratch_in", V), ("scratch_in_bytes", ctypes.c_size_t),
                ("scratch_out", V), ("scratch_out_bytes", ctypes.c_size_t),
                ("deposits", V), ("counts", U32P), ("flags_out", U32P), ("bus_out", U32P),
                ("idx_a", U32P), ("idx_b", U32P), ("idx_c", U32P),
                ("idx_a_src", ctypes.c_size_t), ("idx_b_src", ctypes.c_size_t),
                ("idx_c_src", ctypes.c_size_t), ("idx_scratch_in", U32P),
                ("idx_scratch_src", ctypes.c_size_t), ("lane_mask", V),
                ("lane_mask_bytes", ctypes.c_size_t), ("lane_flags", V),
                ("lane_flags_bytes", ctypes.c_size_t)]


class OpenArgs(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_size_t), ("artifact", ctypes.c_char_p),
                ("index", ctypes.c_int), ("scratch_depth", ctypes.c_uint32)]


def ref_device(ref, depth=256):
    ref.cft_open_ex.argtypes = [ctypes.POINTER(OpenArgs), ctypes.POINTER(V)]
    h = V()
    args = OpenArgs(ctypes.sizeof(OpenArgs), None, 0, depth)
    assert ref.cft_open_ex(ctypes.byref(args), ctypes.byref(h)) == 0
    return h


def ref_load(ref, dev, image):
    ref.cft_program_load.argtypes = [V, V, ctypes.c_size_t, ctypes.POINTER(V)]
    ref.cft_program_free.argtypes = [V]
    img = np.frombuffer(bytes(image), dtype=np.uint8).copy()
    h = V()
    st = ref.cft_program_load(dev, img.ctypes.data if img.nbytes else None, img.nbytes,
                              ctypes.byref(h))
    return st, h


def run_ref(ref, case, depth=256):
    dev = ref_device(ref, depth)
    st, prog = ref_load(ref, dev, case["image"])
    assert st == 0, f"the reference refused the image: {st}"
    out = _outputs(case)
    fl, bus = ctypes.c_uint32(), ctypes.c_uint32()
    A = RunArgs()
    A.struct_size = ctypes.sizeof(A)
    A.a, A.b, A.c, A.n = _p(case["a"]), _p(case["b"]), _p(case["c"]), case["n"]
    if case["bank"] is not None:
        A.bank, A.bank_bytes = _p(case["bank"]), case["bank"].nbytes
    if case["scratch_in"] is not None:
        A.scratch_in, A.scratch_in_bytes = _p(case["scratch_in"]), case["scratch_in"].nbytes
    if out["scratch_out"] is not None:
        A.scratch_out, A.scratch_out_bytes = _p(out["scratch_out"]), out["scratch_out"].nbytes
    A.deposits = _p(out["deposits"])
    A.counts = out["counts"].ctypes.data_as(U32P)
    A.flags_out = ctypes.pointer(fl)
    A.bus_out = ctypes.pointer(bus)
    if case["mask"] is not None:
        A.lane_mask, A.lane_mask_bytes = _p(case["mask"]), case["mask"].nbytes
    A.lane_flags, A.lane_flags_bytes = _p(out["lane_flags"]), case["n"]
    ref.cft_program_run_ex.argtypes = [V, ctypes.POINTER(RunArgs)]
    rc = ref.cft_program_run_ex(prog, ctypes.byref(A))
    ref.cft_program_free(prog)
    ref.cft_close.argtypes = [V]
    ref.cft_close(dev)
    assert rc == 0, f"the reference's run failed: {rc}"
    out["flags"], out["status"] = fl.value, bus.value
    return out


def run_gpu(device, case, prog=None):
    from cftamd import Program
    own = prog is None
    if own:
        prog = Program(device, case["image"])
    out = _outputs(case)
    fl, st = prog.run(case["n"], case["a"], case["b"], case["c"], bank=case["bank"],
                      scratch_in=case["scratch_in"], scratch_out=out["scratch_out"],
                      deposits=out["deposits"], counts=out["counts"],
                      lane_mask=case["mask"], lane_flags=out["lane_flags"])
    if own:
        prog.free()
    out["flags"], out["status"] = fl, st
    return out


def diff(x, y):
    """The outputs on which two runs differ ([] when bit-identical)."""
    bad = []
    for k in ("flags", "status"):
        if x[k] != y[k]:
            bad.append(f"{k} {x[k]:#x} != {y[k]:#x}")
    for k in ("deposits", "counts", "lane_flags", "scratch_out"):
        if x[k] is None and y[k] is None:
            continue
        if not np.array_equal(x[k], y[k]):
            idx = np.flatnonzero(x[k] != y[k])
            bad.append(f"{k} differs at {idx.size} places, first {idx[:4].tolist()}")
    return bad


def oracle_batch(oracle, img, bank, state, steps, fmt, depth=256, threads=16):
    """The batched runner's definition on the oracle: each lane's state
    through its chain of segments - full segments at the image's own count,
    then the remainder - one oracle run per distinct trip count a round.
    Returns (state, flags, status)."""
    esz = ESZ[fmt]
    n = steps.size
    slots = state.nbytes // (n * esz)
    seg = segment_trips(img)
    st = state.reshape(n, slots * esz).copy()
    left = steps.astype(np.int64).copy()
    fl = stat = 0
    while (left > 0).any():
        trips = np.minimum(left, seg)
        for t in sorted(set(trips[left > 0].tolist())):
            ids = np.flatnonzero((left > 0) & (trips == t))
            case = {"image": P.with_trips(img, int(t)), "bank": bank,
                    "a": np.zeros(ids.size * esz, dtype=np.uint8), "b": None, "c": None,
                    "n": ids.size, "fmt": fmt, "scratch_in": st[ids].reshape(-1).copy(),
                    "n_out": slots, "max_dep": 0, "mask": None}
            o = run_oracle(oracle, case, depth=depth, threads=threads)
            st[ids] = o["scratch_out"].reshape(ids.size, slots * esz)
            fl |= o["flags"]
            stat |= o["status"]
        left -= trips
    return st.reshape(-1), fl, stat


segment_trips = P.segment_trips
