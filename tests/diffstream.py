"""Counter-based differential stream on the GPU.

Hard cases of fourteen families are generated on the device by
lib/libcftprobe.so's ``cftp_hardcases`` (cft-fp256_amd/csrc/hardcase.hpp:
case g of a seed is a pure function of the seed, g, the format and the op),
run through libcft.so's ``cft_run`` / ``cft_div`` / ``cft_sqrt``, and
compared with

* the reference's own CPU library (oracle/_ref/libcft_ref.so, compiled from
  its sources): every element's bits, and each call's flag OR;
* the oracle (oracle/sf_oracle.c): every element's bits and every element's
  own flags.

The reference held its fp256 FMA to a stream of the same kind, 2^30
counter-generated hard cases with 0 differing (host/tools/fx256_stream.h:
11-41, docs/studies/GPU-A-hip.md:112-117); the families here aim at the same
classes, with a generator of this library's own.

Per-element flags: an elementwise call returns the OR of its elements'
flags, so the elements are grouped by the flags the oracle gives each of
them and every group is run again on the GPU, in one call of its own: the
group's OR must be exactly its value (each element's flags within it, the
group together raising all of it), and the regrouped results must be the
bits of the first run.

Test infrastructure (it runs the oracle and the reference's CPU library as
the checkers).  ``tests/test_diffstream_gpu.py`` runs short streams; the
command line runs long ones and prints a line a chunk:

    python tests/diffstream.py --fmt 3 --op fma --log2n 30 --chunk-log2 24
    python tests/diffstream.py --path seq --fmt 3 --op fma_negc --log2n 26

With --path seq the same cases go through the sequencer instead: a program
of one operation (or an ABS of a / a NEG of c ahead of an FMA, which the wide
scheduler folds into the FMA's operand as a sign modifier), one trajectory of
one trip a case, through the per-lane runner and the wide runner in turn;
each lane's result and its own flags (lane_flags) against the oracle.
"""

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cft-fp256_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

PROBE_SO = os.path.join(PKG, "lib", "libcftprobe.so")
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libcftoracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libcft_ref.so")

OPS = {"fma": 0, "add": 1, "sub": 2, "mul": 3, "div": 256, "sqrt": 257}
OP_NAMES = {v: k for k, v in OPS.items()}
FAMILIES = ["raw bits", "subnormal-edge operands", "product at underflow",
            "product at overflow", "deep cancellation", "exact cancellation / product error",
            "ties of a*b + c", "c over +-(2P+40) binades", "specials", "cancellation near EMIN",
            "cancellation near EMAX", "c at the normal boundary", "overflow by a rounding carry",
            "alignment edges"]
NFAM = len(FAMILIES)
FLAG_BITS = ["invalid", "divbyzero", "overflow", "underflow", "inexact"]
V = ctypes.c_void_p
U32P = ctypes.POINTER(ctypes.c_uint32)


def reads_b(op):
    return op in (0, 3, 256)


def reads_c(op):
    return op in (0, 1, 2)


def esize(fmt):
    return 4 << fmt


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def load_probe():
    P = ctypes.CDLL(PROBE_SO)
    P.cftp_hardcases.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                 ctypes.c_size_t, V, V, V, V]
    return P


def load_oracle():
    O = ctypes.CDLL(ORACLE_SO)
    O.oc_run_lanes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                               ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P, V]
    return O


def load_ref():
    if not os.path.exists(REF_SO):
        return None
    R = ctypes.CDLL(REF_SO)
    R.cft_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(V)]
    R.cft_close.argtypes = [V]
    R.cft_run.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                          ctypes.c_size_t, U32P, U32P]
    R.cft_div.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, V, ctypes.c_size_t, U32P, U32P]
    R.cft_sqrt.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, ctypes.c_size_t, U32P, U32P]
    return R


def _p(x):
    return None if x is None else x.ctypes.data


def generate(P, fmt, op, seed, base, n):
    """Cases base .. base+n-1 generated on the GPU and copied back: (a, b, c,
    family), a/b/c as (n, esz) uint8 arrays (None where op reads none)."""
    from cftamd import DeviceArray
    es = esize(fmt)
    da = DeviceArray(n * es)
    db = DeviceArray(n * es) if reads_b(op) else None
    dc = DeviceArray(n * es) if reads_c(op) else None
    df = DeviceArray(n)
    try:
        st = P.cftp_hardcases(fmt, op, seed, base, n, da.ptr, db.ptr if db else None,
                              dc.ptr if dc else None, df.ptr)
        if st != 0:
            raise RuntimeError(f"cftp_hardcases returned {st}")
        out = [x.download().reshape(n, es) if x is not None else None for x in (da, db, dc)]
        fam = df.download()
    finally:
        for x in (da, db, dc, df):
            if x is not None:
                x.free()
    return out[0], out[1], out[2], fam


def gpu_call(dev, fmt, op, rnd, a, b, c):
    if op == 256:
        d, fl = dev.div(fmt, rnd, a, b)
    elif op == 257:
        d, fl = dev.sqrt(fmt, rnd, a)
    else:
        d, fl = dev.run(op, fmt, rnd, a, b if reads_b(op) else None, c if reads_c(op) else None)
    return d.reshape(-1, esize(fmt)), fl


def ref_call(R, fmt, op, rnd, a, b, c, threads):
    """The reference's CPU library over slices, one software device a
    thread; returns (d, flag OR)."""
    n, es = a.shape
    d = np.empty((n, es), np.uint8)
    flags, bad = [0] * threads, []

    def work(t):
        lo, hi = n * t // threads, n * (t + 1) // threads
        if hi <= lo:
            return
        dev = V()
        if R.cft_open(None, 0, ctypes.byref(dev)) != 0:
            bad.append("cft_open")
            return
        f = ctypes.c_uint32()
        sl = lambda x: None if x is None else x[lo:hi].ctypes.data
        if op == 256:
            st = R.cft_div(dev, fmt, rnd, sl(a), sl(b), sl(d), hi - lo, ctypes.byref(f), None)
        elif op == 257:
            st = R.cft_sqrt(dev, fmt, rnd, sl(a), sl(d), hi - lo, ctypes.byref(f), None)
        else:
            st = R.cft_run(dev, op, fmt, rnd, sl(a), sl(b) if reads_b(op) else None,
                           sl(c) if reads_c(op) else None, sl(d), hi - lo, ctypes.byref(f), None)
        R.cft_close(dev)
        if st != 0:
            bad.append(st)
        flags[t] = f.value

    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    if bad:
        raise RuntimeError(f"the reference refused a call: {bad[0]}")
    fl = 0
    for f in flags:
        fl |= f
    return d, fl


def oracle_call(O, fmt, op, rnd, a, b, c, threads):
    """The oracle: (d, flag OR, each element's flags)."""
    n, es = a.shape
    d = np.empty((n, es), np.uint8)
    lanes = np.empty(n, np.uint8)
    fl = ctypes.c_uint32()
    O.oc_run_lanes(fmt, op, rnd, _p(a), _p(b) if reads_b(op) else None,
                   _p(c) if reads_c(op) else None, _p(d), n, 0, threads, ctypes.byref(fl),
                   _p(lanes))
    return d, fl.value, lanes


def check_chunk(dev, R, O, fmt, op, rnd, a, b, c, fam, threads):
    """One call's differential: returns a dict of counts (and the indices of
    the first differing elements)."""
    n = a.shape[0]
    got, gfl = gpu_call(dev, fmt, op, rnd, a, b, c)
    want_o, ofl, lanes = oracle_call(O, fmt, op, rnd, a, b, c, threads)
    diff_o = (got != want_o).any(axis=1)
    res = {"cases": n, "gpu_flags": gfl, "oracle_flags": ofl}
    if R is not None:
        want_r, rfl = ref_call(R, fmt, op, rnd, a, b, c, threads)
        diff_r = (got != want_r).any(axis=1)
        res["ref_flags"] = rfl
        res["oracle_vs_ref"] = int(np.count_nonzero((want_o != want_r).any(axis=1)))
    else:
        diff_r = np.zeros(n, bool)
    # per-element flags: one GPU call a group of equal oracle flags
    order = np.argsort(lanes, kind="stable")
    sl = lanes[order]
    vals, starts = np.unique(sl, return_index=True)
    ends = list(starts[1:]) + [n]
    ga = a[order]
    gb = b[order] if reads_b(op) else None
    gc = c[order] if reads_c(op) else None
    group_bad, regroup_bits = [], 0
    for v, s, e in zip(vals, starts, ends):
        d2, f2 = gpu_call(dev, fmt, op, rnd, ga[s:e], None if gb is None else gb[s:e],
                          None if gc is None else gc[s:e])
        if f2 != int(v):
            group_bad.append((int(v), int(f2), int(e - s)))
        regroup_bits += int(np.count_nonzero((d2 != got[order[s:e]]).any(axis=1)))
    res["differ_ref"] = int(np.count_nonzero(diff_r))
    res["differ_oracle"] = int(np.count_nonzero(diff_o))
    res["flag_or_ok"] = (gfl == ofl) and (R is None or gfl == res["ref_flags"])
    res["flag_groups"] = int(vals.size)
    res["flag_groups_bad"] = group_bad
    res["regroup_bits_differ"] = regroup_bits
    bad = np.flatnonzero(diff_r | diff_o)
    res["first_bad"] = [int(i) for i in bad[:8]]
    # per family: cases, differing, and each flag bit's count (the oracle's)
    fam_cases = np.bincount(fam, minlength=NFAM)
    fam_bad = np.bincount(fam[diff_r | diff_o], minlength=NFAM)
    fam_flags = np.stack([np.bincount(fam, weights=(lanes >> k) & 1, minlength=NFAM)
                          for k in range(5)], axis=1).astype(np.int64)
    res["fam_cases"] = fam_cases.tolist()
    res["fam_bad"] = fam_bad.tolist()
    res["fam_flags"] = fam_flags.tolist()
    return res


def stream(fmt, op, total, chunk, seed=1, modes=(0, 1, 2, 3, 4), all_modes=False, dev=None,
           threads=None, log=print, with_ref=True):
    """`total` cases in chunks of `chunk`; chunk j under mode modes[j % len]
    (every mode where all_modes).  Returns the summary dict."""
    from cftamd import Device
    threads = threads or cpu_threads()
    P, O = load_probe(), load_oracle()
    R = load_ref() if with_ref else None
    own = dev is None
    if own:
        dev = Device("hip:0")
    S = {"format": fmt, "op": OP_NAMES[op], "seed": seed, "cases": 0, "calls": 0,
         "differ_ref": 0, "differ_oracle": 0, "oracle_vs_ref": 0, "flag_or_bad": 0,
         "flag_groups_bad": 0, "regroup_bits_differ": 0, "reference": R is not None,
         "per_mode_cases": [0] * 5, "fam_cases": [0] * NFAM, "fam_bad": [0] * NFAM,
         "fam_flags": [[0] * 5 for _ in range(NFAM)], "first_bad": [], "threads": threads}
    t0 = time.perf_counter()
    try:
        nchunks = (total + chunk - 1) // chunk
        for j in range(nchunks):
            base = j * chunk
            n = min(chunk, total - base)
            a, b, c, fam = generate(P, fmt, op, seed, base, n)
            for rnd in (modes if all_modes else (modes[j % len(modes)],)):
                r = check_chunk(dev, R, O, fmt, op, rnd, a, b, c, fam, threads)
                S["cases"] += n
                S["calls"] += 1
                S["per_mode_cases"][rnd] += n
                for k in ("differ_ref", "differ_oracle", "regroup_bits_differ"):
                    S[k] += r[k]
                S["oracle_vs_ref"] += r.get("oracle_vs_ref", 0)
                S["flag_or_bad"] += 0 if r["flag_or_ok"] else 1
                S["flag_groups_bad"] += len(r["flag_groups_bad"])
                for f in range(NFAM):
                    S["fam_cases"][f] += r["fam_cases"][f]
                    S["fam_bad"][f] += r["fam_bad"][f]
                    for k in range(5):
                        S["fam_flags"][f][k] += r["fam_flags"][f][k]
                if r["first_bad"] and len(S["first_bad"]) < 8:
                    S["first_bad"] += [(base + i, rnd) for i in r["first_bad"]]
                log(f"diffstream fmt {fmt} {OP_NAMES[op]} chunk {j + 1}/{nchunks} base {base} "
                    f"n {n} rnd {rnd}: differ ref {r['differ_ref']} oracle "
                    f"{r['differ_oracle']}, flag OR {'ok' if r['flag_or_ok'] else 'BAD'}, "
                    f"{r['flag_groups']} flag groups, {len(r['flag_groups_bad'])} bad, "
                    f"{time.perf_counter() - t0:.1f} s")
    finally:
        if own:
            dev.close()
    S["seconds"] = round(time.perf_counter() - t0, 1)
    S["clean"] = (S["differ_ref"] == 0 and S["differ_oracle"] == 0 and S["flag_or_bad"] == 0
                  and S["flag_groups_bad"] == 0 and S["regroup_bits_differ"] == 0
                  and S["oracle_vs_ref"] == 0)
    return S


# ---- the sequencer on the same cases ----------------------------------------

# program kind -> (the generator's op for its operands, the oracle's op, the
# operand changes the program makes before it: a fused ABS of a, NEG of c)
SEQ_KINDS = {"fma": (0, 0, None), "fma_negc": (0, 0, "negc"), "fma_absa": (0, 0, "absa"),
             "add": (1, 1, None), "sub": (2, 2, None), "mul": (3, 3, None)}


def seq_program(kind, rnd, fmt):
    """slots 0..2 -> r1..r3, one operation (with an ABS of a or a NEG of c
    ahead of the FMA for the fused kinds), the result to slot 0."""
    from cftamd import programs as P
    ins = [P.ldl(1, 0), P.ldl(2, 1), P.ldl(3, 2), P.repeat(1)]
    if kind == "fma":
        ins.append(P.alu(P.FMA, 4, 1, 2, 3, rnd=rnd))
    elif kind == "fma_negc":
        ins += [P.alu(P.NEG, 5, 3), P.alu(P.FMA, 4, 1, 2, 5, rnd=rnd)]
    elif kind == "fma_absa":
        ins += [P.alu(P.ABS, 5, 1), P.alu(P.FMA, 4, 5, 2, 3, rnd=rnd)]
    elif kind == "add":
        ins.append(P.alu(P.ADD, 4, 1, 0, 3, rnd=rnd))
    elif kind == "sub":
        ins.append(P.alu(P.SUB, 4, 1, 0, 3, rnd=rnd))
    else:
        ins.append(P.alu(P.MUL, 4, 1, 2, rnd=rnd))
    ins += [P.endrep(), P.stl(4, 0), P.halt()]
    return P.image(fmt, ins, scratch_in=3, scratch_out=3)


def seq_check_chunk(dev, O, fmt, kind, rnd, a, b, c, fam, width, threads):
    """One batched run of the kind's program over the chunk's cases, one
    trajectory of one trip a case, at `width` lanes an instance (1: the
    per-lane interpreter; 2 or more: the wide runner): every lane's result
    and its own flags against the oracle."""
    from cftamd import Program
    gop, oop, mod = SEQ_KINDS[kind]
    n, es = a.shape
    zero = np.zeros((n, es), np.uint8)
    xb = b if b is not None else zero
    xc = c if c is not None else zero
    state = np.ascontiguousarray(np.stack([a, xb, xc], axis=1)).reshape(-1)
    lf = np.zeros(n, np.uint8)
    prog = Program(dev, seq_program(kind, rnd, fmt))
    try:
        gfl, status = prog.run_batch(state, np.ones(n, np.uint32), lane_flags=lf,
                                     lanes_per_instance=width)
    finally:
        prog.free()
    got = state.reshape(n, 3, es)[:, 0, :]
    oa, oc = a, xc
    if mod == "absa":
        oa = a.copy()
        oa[:, -1] &= 0x7F
    elif mod == "negc":
        oc = xc.copy()
        oc[:, -1] ^= 0x80
    want, ofl, lanes = oracle_call(O, fmt, oop, rnd, oa, xb, oc, threads)
    diff = (got != want).any(axis=1)
    fdiff = (lf & 0x1F) != lanes
    bad = diff | fdiff
    return {"cases": n, "differ": int(np.count_nonzero(diff)),
            "flags_differ": int(np.count_nonzero(fdiff)), "status": int(status),
            "flag_or_ok": gfl == ofl, "first_bad": [int(i) for i in np.flatnonzero(bad)[:8]],
            "fam_bad": np.bincount(fam[bad], minlength=NFAM).tolist()}


def seq_stream(fmt, kind, total, chunk, seed=1, modes=(0, 1, 2, 3, 4), widths=(1, 2), dev=None,
               threads=None, log=print):
    """The sequencer's batched runner on `total` hard cases of the kind's
    generator stream, chunk j under mode modes[j % len] at width
    widths[j % len] (so both runners see every mode over enough chunks)."""
    from cftamd import Device
    threads = threads or cpu_threads()
    P_, O = load_probe(), load_oracle()
    gop = SEQ_KINDS[kind][0]
    own = dev is None
    if own:
        dev = Device("hip:0")
    S = {"format": fmt, "kind": kind, "seed": seed, "cases": 0, "differ": 0, "flags_differ": 0,
         "flag_or_bad": 0, "status_bad": 0, "by_width": {}, "per_mode_cases": [0] * 5,
         "fam_bad": [0] * NFAM, "first_bad": [], "threads": threads}
    t0 = time.perf_counter()
    try:
        nchunks = (total + chunk - 1) // chunk
        for j in range(nchunks):
            base = j * chunk
            n = min(chunk, total - base)
            a, b, c, fam = generate(P_, fmt, gop, seed, base, n)
            rnd = modes[j % len(modes)]
            w = widths[(j // len(modes)) % len(widths)] if nchunks > len(modes) else \
                widths[j % len(widths)]
            r = seq_check_chunk(dev, O, fmt, kind, rnd, a, b, c, fam, w, threads)
            S["cases"] += n
            S["per_mode_cases"][rnd] += n
            S["by_width"][w] = S["by_width"].get(w, 0) + n
            S["differ"] += r["differ"]
            S["flags_differ"] += r["flags_differ"]
            S["flag_or_bad"] += 0 if r["flag_or_ok"] else 1
            S["status_bad"] += 0 if r["status"] == 0 else 1
            for f in range(NFAM):
                S["fam_bad"][f] += r["fam_bad"][f]
            if r["first_bad"] and len(S["first_bad"]) < 8:
                S["first_bad"] += [(base + i, rnd, w) for i in r["first_bad"]]
            log(f"diffstream seq fmt {fmt} {kind} chunk {j + 1}/{nchunks} base {base} n {n} "
                f"rnd {rnd} width {w}: differ {r['differ']}, flags differ {r['flags_differ']}, "
                f"{time.perf_counter() - t0:.1f} s")
    finally:
        if own:
            dev.close()
    S["seconds"] = round(time.perf_counter() - t0, 1)
    S["clean"] = (S["differ"] == 0 and S["flags_differ"] == 0 and S["flag_or_bad"] == 0 and
                  S["status_bad"] == 0)
    return S


def table(S):
    """A per-family table of one stream's summary (markdown)."""
    lines = [f"| family | cases | differing | " + " | ".join(FLAG_BITS) + " |",
             "|---|---:|---:|" + "---:|" * 5]
    for f in range(NFAM):
        fl = " | ".join(str(x) for x in S["fam_flags"][f])
        lines.append(f"| {f} {FAMILIES[f]} | {S['fam_cases'][f]} | {S['fam_bad'][f]} | {fl} |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--fmt", type=int, default=3)
    ap.add_argument("--op", default="fma", choices=sorted(OPS) + sorted(k for k in SEQ_KINDS
                                                                      if k not in OPS))
    ap.add_argument("--path", default="elem", choices=["elem", "seq"],
                    help="elem: cft_run/cft_div/cft_sqrt; seq: a one-operation program "
                         "through the batched runners (per-lane and wide)")
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--chunk-log2", type=int, default=22)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--all-modes", action="store_true",
                    help="every chunk under all five modes (default: one mode a chunk)")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--json", default="", help="append the summary to this file")
    args = ap.parse_args()
    if args.path == "seq":
        S = seq_stream(args.fmt, args.op, 1 << args.log2n,
                       1 << min(args.chunk_log2, args.log2n), seed=args.seed,
                       threads=args.threads or None, log=lambda s: print(s, flush=True))
        print(json.dumps(S), flush=True)
        if args.json:
            with open(args.json, "a") as f:
                f.write(json.dumps(S) + "\n")
        sys.exit(0 if S["clean"] else 1)
    S = stream(args.fmt, OPS[args.op], 1 << args.log2n, 1 << min(args.chunk_log2, args.log2n),
               seed=args.seed, all_modes=args.all_modes, threads=args.threads or None,
               log=lambda s: print(s, flush=True))
    print(table(S), flush=True)
    out = {k: v for k, v in S.items()}
    print(json.dumps(out), flush=True)
    if args.json:
        with open(args.json, "a") as f:
            f.write(json.dumps(out) + "\n")
    sys.exit(0 if S["clean"] else 1)


if __name__ == "__main__":
    main()
