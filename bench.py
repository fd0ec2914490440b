#!/usr/bin/env python3
"""Benchmark of the MI355X engine on the reference's headline workloads.

Headline (value): binary256 FMA over the elementwise workload - N = 2^24
elements from seed 0, exponents over the full range, 3% subnormals, 2%
specials, 5% FMA cases constructed near a rounding boundary.  One step is the
FMA over all N elements in each of the five rounding modes (five cft_run
calls), operands resident in HBM (cft_alloc buffers) before the timed region.

Beside it, in the same JSON line:
  ops         binary256 add, mul, div and sqrt over the same operands;
  rk4         Lorenz-63 RK4 in binary256, 65,536 trajectories of
              log-uniform 1e3..1e5 steps, h = RN(1e-3): the batched runner in
              one persistent launch through a HIP graph (steps/s), the plain
              one-launch-per-instruction form on a 1/100 sample of the same
              ensemble beside the persistent form on that sample (bit-equal),
              and the reference's CPU program runs on a bounded sample;
  lorenz96    Lorenz-96 RK4 in binary64 and binary256, 8,192 instances, ring
              sizes 8..256, steps uniform 1e3..2e4: every ring size in one
              multi-program launch, and on a 1/100 sample of the steps that
              launch beside one batched launch a ring size (bit-equal);
  serve       the serving loop over a Poisson stream at 2,000 requests/s:
              p50/p99 latency, throughput, every job checked against the job
              run on its own, the plain one-job-a-launch form on a prefix;
  cpu_baseline  the reference's CPU FMA on this host's cores, same inputs,
              compared bit for bit, all threads and one core, with the host's
              CPU model and counts;
  roofline    the FMA kernel against 8 TB/s, and against the rate a library
              ternary elementwise kernel (torch.addcmul) streams over the same
              bytes (stream_GBs, frac_of_stream); beside it the VALU figures:
              the issue rates of the integer instructions measured live
              (lib/libcftprobe.so), the FMA's VALU issue and its algorithmic
              VALU fraction (64 multiply-accumulates an element at their
              measured issue cost).

Checks in the same run, each reported with its size: add and mul over all
5 x 2^24 results against the reference's CPU library, div and sqrt on every
16th element of each mode; the Lorenz-63 ensemble's final states of its 32
longest and 224 rank-stratified trajectories against the reference's chain
of segments; one Lorenz-96 instance a ring size in each format; every
served elementwise job's bits against the job run alone, 32 elements of each
(and every element of jobs up to 20,000) against the reference, and a
stratified sample of the served ODE lanes against the reference.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --copy-calibration BYTES     (a plain device copy, for
                                                  scaling HBM counters)
"""

import argparse
import ctypes
import json
from concurrent.futures import ThreadPoolExecutor
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cft-fp256_amd"))

METRIC = "bit-exact binary256 FMA Gelem/s; batched fp256 RK4 steps/s; p99 serve latency"
HBM_PEAK_GBS = 8000.0
FMA_BYTES = 128          # 3 x 32 B in, 32 B out per element
# HBM bytes an FMA element actually moves: rocprofv3 FETCH_SIZE (x2.000) +
# WRITE_SIZE (x1.000) at 2^22 elements, each scaled by the same counter on a
# plain 256 MiB copy (profiles/fma_fp256_elementwise.md)
FMA_TRAFFIC_PER_ELEM = (393510.0 + 131151.0) * 1024 / 4194304
BYTES = {"add": 96, "mul": 96, "fma": 128, "div": 96, "sqrt": 64}
# VALU: 256 CUs x 4 SIMDs x 2.4 GHz.  gfx950 issues a plain 32-bit wave64
# operation every 2 cycles (32 lanes a cycle: the 157.3 TFLOP/s vector fp32
# rate) and a carry, multiply or v_mad_u64_u32 every 4 (the probes measure
# 2.9 and 4.3-4.7 cycles with loop control, profiles/valu_issue_rates.md)
VALU_PEAK_WAVE_INSNS = 256 * 4 * 2.4e9 / 2
VALU_4CYCLE_WAVE_INSNS = 256 * 4 * 2.4e9 / 4
# VALU instructions a lane issues for one binary256 FMA element of this
# workload (rocprofv3 SQ_INSTS_VALU / SQ_WAVES at 2^22 elements, and the
# cycles a SIMD spends a VALU instruction there, GRBM_GUI_ACTIVE / 8 XCDs
# over SQ_INSTS_VALU / 1,024 SIMDs: profiles/valu_issue_rates.md); the exact
# product's 32x32->64 multiply-accumulates an element (8 x 8 limbs)
FMA_VALU_PER_ELEM = 640.8
FMA_CYCLES_PER_VALU = 5.65
FMA_MACS = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-ops", action="store_true", help="FMA only")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-rk4", action="store_true", help="no Lorenz-63/96 sections")
    ap.add_argument("--skip-serve", action="store_true", help="no serving section")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES, three times")
    return ap.parse_args()


def host_info():
    """nproc, the CPUs this process may use, physical cores and the CPU
    model, from the kernel's own tables."""
    model, phys = None, set()
    try:
        with open("/proc/cpuinfo") as fh:
            pid = None
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    phys.add((pid, v))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "cpus_available": avail,
            "physical_cores": len(phys) or None, "cpu_model": model}


def cpu_threads(args, info):
    """Threads for the CPU legs: --cpu-threads, else one per physical core
    the process may use (OMP_NUM_THREADS, where set, is the host's share)."""
    if args.cpu_threads:
        return args.cpu_threads
    t = min(info["cpus_available"], info["physical_cores"] or info["cpus_available"])
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(t, omp) if omp > 0 else t)


class RefCPU:
    """The reference's own C library (oracle/_ref/libcft_ref.so, compiled
    from its sources by oracle/Makefile): its software device, one per
    thread over contiguous slices.  `ok` is False where it is not built."""

    def __init__(self):
        path = os.path.join(HERE, "oracle", "_ref", "libcft_ref.so")
        self.ok = os.path.exists(path)
        if not self.ok:
            return
        L = ctypes.CDLL(path)
        V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)
        L.cft_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(V)]
        L.cft_close.argtypes = [V]
        L.cft_run.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                              ctypes.c_size_t, U32P, U32P]
        L.cft_div.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, V, ctypes.c_size_t, U32P,
                              U32P]
        L.cft_sqrt.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V, ctypes.c_size_t, U32P, U32P]
        self.L = L

    def elem(self, name, fmt, rnd, a, b, c, n, threads):
        """d = op(a, b, c) for name in fma/add/mul/div/sqrt (ADD reads a and
        c, MUL a and b); returns (d, flags ORed, seconds)."""
        L, V = self.L, ctypes.c_void_p
        esz = {0: 4, 1: 8, 2: 16, 3: 32}[fmt]
        d = np.empty(n * esz, dtype=np.uint8)
        flags = [0] * threads
        bad = []
        p = lambda x, off: None if x is None else x.ctypes.data + off

        def work(t):
            lo, hi = n * t // threads, n * (t + 1) // threads
            if hi <= lo:
                return
            dev = V()
            L.cft_open(None, 0, ctypes.byref(dev))
            f, o = ctypes.c_uint32(), esz * lo
            if name == "fma":
                st = L.cft_run(dev, 0, fmt, rnd, p(a, o), p(b, o), p(c, o), p(d, o), hi - lo,
                               ctypes.byref(f), None)
            elif name == "add":
                st = L.cft_run(dev, 1, fmt, rnd, p(a, o), None, p(c, o), p(d, o), hi - lo,
                               ctypes.byref(f), None)
            elif name == "mul":
                st = L.cft_run(dev, 3, fmt, rnd, p(a, o), p(b, o), None, p(d, o), hi - lo,
                               ctypes.byref(f), None)
            elif name == "div":
                st = L.cft_div(dev, fmt, rnd, p(a, o), p(b, o), p(d, o), hi - lo,
                               ctypes.byref(f), None)
            else:
                st = L.cft_sqrt(dev, fmt, rnd, p(a, o), p(d, o), hi - lo, ctypes.byref(f), None)
            L.cft_close(dev)
            if st != 0:
                bad.append(st)
            flags[t] = f.value
        t0 = time.perf_counter()
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
        return d, fl, time.perf_counter() - t0


def issue_rates():
    """Wave instructions a second the GPU issues of v_add_u32, v_mad_u64_u32,
    v_mul_lo_u32, v_mul_hi_u32 and a v_add_co_u32 + v_addc_co_u32 pair
    (lib/libcftprobe.so: 8 independent chains a lane, 8 waves a SIMD), and
    each relative to v_add_u32."""
    path = os.path.join(HERE, "cft-fp256_amd", "lib", "libcftprobe.so")
    if not os.path.exists(path):
        return None
    P = ctypes.CDLL(path)
    P.cftp_issue_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
    out = {}
    for kind, name in enumerate(("v_add_u32", "v_mad_u64_u32", "v_mul_lo_u32",
                                 "v_mul_hi_u32", "v_add_co_u32+v_addc_co_u32")):
        ms, wi = ctypes.c_double(), ctypes.c_double()
        st = P.cftp_issue_rate(kind, 2048, 16384, 5, ctypes.byref(ms), ctypes.byref(wi))
        if st != 0:
            return None
        out[name] = wi.value / (ms.value * 1e-3)
    rel = {k: round(v / out["v_add_u32"], 4) for k, v in out.items()}
    return {"wave_insns_per_s": {k: round(v / 1e9, 1) for k, v in out.items()},
            "unit": "G wave instructions/s", "relative_to_v_add_u32": rel,
            "probe": "lib/libcftprobe.so cftp_issue_rate: 2,048 workgroups x 256 lanes, "
                     "16,384 rounds of 8 independent chains, mean of 5 launches"}


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size (three times), the
    reference point for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"bench: copied {nbytes} bytes three times")


def stream_probe(n_elem, elem_bytes=32, reps=5):
    """The streaming rate a library ternary elementwise kernel reaches on the
    FMA's byte pattern: torch.addcmul over float32 tensors of n_elem *
    elem_bytes bytes each, three read and one written, timed with events on
    torch's stream.  The practical ceiling beside the 8 TB/s peak."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
        k = n_elem * elem_bytes // 4
        a, b, c = (torch.rand(k, device="cuda") for _ in range(3))
        d = torch.empty_like(a)
        torch.addcmul(c, a, b, out=d)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            torch.addcmul(c, a, b, out=d)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del a, b, c, d
        torch.cuda.empty_cache()
        return 4 * n_elem * elem_bytes / (ms * 1e-3) / 1e9
    except Exception as e:                       # a measurement aid, never the result
        log(f"bench: stream probe skipped ({e})")
        return None


def ref_program_chain(img, bank, state, steps, fmt, threads, depth=2048):
    """The reference's CPU program runs (oracle/_ref, its own C library) on a
    chain of segments per lane - full segments at the image's trip count,
    then the remainder - one software device per thread over slices of
    lanes, each opened at scratch depth `depth` (cft_open_ex; Lorenz-96 at
    ring K needs 5K slots).  Returns (final states, flags, status, seconds)
    or None."""
    import ctypes as C
    sys.path.insert(0, os.path.join(HERE, "cft-fp256_amd"))
    from cftamd import programs as P
    path = os.path.join(HERE, "oracle", "_ref", "libcft_ref.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    V, U32P = C.c_void_p, C.POINTER(C.c_uint32)

    class RunArgs(C.Structure):
        _fields_ = [("struct_size", C.c_size_t), ("a", V), ("b", V), ("c", V),
                    ("n", C.c_size_t), ("bank", V), ("bank_bytes", C.c_size_t),
                    ("scratch_in", V), ("scratch_in_bytes", C.c_size_t),
                    ("scratch_out", V), ("scratch_out_bytes", C.c_size_t),
                    ("deposits", V), ("counts", U32P), ("flags_out", U32P),
                    ("bus_out", U32P), ("idx_a", U32P), ("idx_b", U32P), ("idx_c", U32P),
                    ("idx_a_src", C.c_size_t), ("idx_b_src", C.c_size_t),
                    ("idx_c_src", C.c_size_t), ("idx_scratch_in", U32P),
                    ("idx_scratch_src", C.c_size_t), ("lane_mask", V),
                    ("lane_mask_bytes", C.c_size_t), ("lane_flags", V),
                    ("lane_flags_bytes", C.c_size_t)]
    class OpenArgs(C.Structure):
        _fields_ = [("struct_size", C.c_size_t), ("artifact", C.c_char_p),
                    ("index", C.c_int), ("scratch_depth", C.c_uint32)]
    lib.cft_open_ex.argtypes = [C.POINTER(OpenArgs), C.POINTER(V)]
    lib.cft_close.argtypes = [V]
    lib.cft_program_load.argtypes = [V, V, C.c_size_t, C.POINTER(V)]
    lib.cft_program_free.argtypes = [V]
    lib.cft_program_run_ex.argtypes = [V, C.POINTER(RunArgs)]
    esz = {1: 8, 2: 16, 3: 32}[fmt]
    n = steps.size
    slots = state.nbytes // (n * esz)
    seg = P.segment_trips(img)
    st = state.reshape(n, slots * esz).copy()
    res = {"flags": 0, "status": 0, "bad": 0}
    lock = threading.Lock()

    def work(lo, hi):
        dev = V()
        oa = OpenArgs(C.sizeof(OpenArgs), None, 0, depth)
        if lib.cft_open_ex(C.byref(oa), C.byref(dev)) != 0:
            res["bad"] += 1
            return
        progs = {}
        left = steps[lo:hi].astype(np.int64).copy()
        fl = stt = 0
        while (left > 0).any():
            trips = np.minimum(left, seg)
            for t in sorted(set(trips[left > 0].tolist())):
                ids = lo + np.flatnonzero((left > 0) & (trips == t))
                if t not in progs:
                    im = np.frombuffer(P.with_trips(img, int(t)), dtype=np.uint8).copy()
                    h = V()
                    if lib.cft_program_load(dev, im.ctypes.data, im.nbytes, C.byref(h)) != 0:
                        res["bad"] += 1
                        return
                    progs[t] = (h, im)
                sin = st[ids].reshape(-1).copy()
                sout = np.empty_like(sin)
                a = np.zeros(ids.size * esz, dtype=np.uint8)
                f, b = C.c_uint32(), C.c_uint32()
                A = RunArgs()
                A.struct_size = C.sizeof(A)
                A.a, A.n = a.ctypes.data, ids.size
                A.bank, A.bank_bytes = bank.ctypes.data, bank.nbytes
                A.scratch_in, A.scratch_in_bytes = sin.ctypes.data, sin.nbytes
                A.scratch_out, A.scratch_out_bytes = sout.ctypes.data, sout.nbytes
                A.flags_out, A.bus_out = C.pointer(f), C.pointer(b)
                if lib.cft_program_run_ex(progs[t][0], C.byref(A)) != 0:
                    res["bad"] += 1
                    return
                st[ids] = sout.reshape(ids.size, slots * esz)
                fl |= f.value
                stt |= b.value
            left -= trips
        for h, _im in progs.values():
            lib.cft_program_free(h)
        lib.cft_close(dev)
        with lock:
            res["flags"] |= fl
            res["status"] |= stt

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(n * t // threads, n * (t + 1) // threads))
           for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    secs = time.perf_counter() - t0
    if res["bad"]:
        return None
    return st.reshape(-1), res["flags"], res["status"], secs


def rk4_section(dev, lib, args, threads, ref):
    """Lorenz-63 RK4 binary256 ensemble: persistent batched runner on the whole
    ensemble, per-op form against it on a 1/100 sample, the reference's CPU
    runs on a bounded sample."""
    from cftamd import Device, Program, odes, _lib as L
    out = {}
    img, bank, state, steps = odes.lorenz63(65536, seed=args.seed, fmt=3)
    total = int(steps.sum())
    prog = Program(dev, img)
    sbuf = dev.alloc(state.nbytes)
    bbuf = dev.alloc(bank.nbytes)
    bbuf.host[:] = bank
    bbuf.to_device()
    # the autotuner on a 1/100 sample of the steps: lanes an instance (1 to
    # 16, or the library's model) and workgroup size, by measurement
    from cftamd import tune_batch
    prog.prepare(0x7F)
    s_tune = np.maximum(1, steps // 100).astype(np.uint32)
    knobs, trials = tune_batch(prog, state, s_tune, bank=bank, repeats=2)
    log(f"bench: rk4 tuned {knobs} ({len(trials)} trials, {trials[0][0]:.4f} .. "
        f"{trials[-1][0]:.4f} s)")
    # warm the graph on this exact launch shape: a pass over a copy of the
    # ensemble with one step a lane, one untimed pass over the real steps (a
    # process's first seconds-long launch runs ~10% slower: the card's
    # clocks), then the timed pass
    sbuf.host[:] = state
    sbuf.to_device()
    prog.run_batch(sbuf, np.ones_like(steps), bank=bbuf, graph=True, **knobs)
    sbuf.from_device()
    sbuf.host[:] = state
    sbuf.to_device()
    t0 = time.perf_counter()
    prog.run_batch(sbuf, steps, bank=bbuf, graph=True, **knobs)
    first_s = time.perf_counter() - t0
    sbuf.from_device()
    sbuf.host[:] = state
    sbuf.to_device()
    lib.cftx_kernel_timer(dev.handle, 1)
    t0 = time.perf_counter()
    fl, st = prog.run_batch(sbuf, steps, bank=bbuf, graph=True, **knobs)
    secs = time.perf_counter() - t0
    kms, kl = ctypes.c_double(), ctypes.c_uint64()
    lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(kl))
    lib.cftx_kernel_timer(dev.handle, 0)
    sbuf.from_device()
    full = sbuf.host.copy()
    out["lorenz63_fp256"] = {
        "trajectories": int(steps.size), "total_steps": total,
        "steps_min": int(steps.min()), "steps_max": int(steps.max()),
        "persistent_s": round(secs, 4), "kernel_s": round(kms.value / 1e3, 4),
        "steps_per_s": round(total / secs, 1), "alu_ops_per_s": round(53 * total / secs, 1),
        "untimed_first_pass_s": round(first_s, 4),
        "flags": fl, "status": st, "launches": int(lib.cftx_batch_launches(dev.handle)),
        "lanes_per_instance": int(lib.cftx_batch_width(dev.handle)),
        "tuned_knobs": knobs, "tuning_sample": "steps // 100 (at least 1)",
        "tuning_trials_s": [[t, k] for t, k in trials]}
    log(f"bench: rk4 persistent {secs:.3f} s -> {total / secs / 1e6:.1f} M steps/s")

    # the plain per-op form on a 1/100 sample, beside the persistent form
    s2 = np.maximum(1, steps // 100).astype(np.uint32)
    ta = dev.alloc(state.nbytes)
    ta.host[:] = state
    ta.to_device()
    t0 = time.perf_counter()
    prog.run_batch(ta, s2, bank=bbuf, graph=True, **knobs)
    tp = time.perf_counter() - t0
    tb = dev.alloc(state.nbytes)
    tb.host[:] = state
    tb.to_device()
    t0 = time.perf_counter()
    prog.run_batch(tb, s2, bank=bbuf, mode=L.BATCH_PER_OP)
    to = time.perf_counter() - t0
    nlaunch = int(lib.cftx_batch_launches(dev.handle))
    ta.from_device()
    tb.from_device()
    out["lorenz63_fp256_sample"] = {
        "sample": "the same ensemble at steps // 100 (at least 1)",
        "total_steps": int(s2.sum()), "persistent_s": round(tp, 4), "per_op_s": round(to, 4),
        "per_op_launches": nlaunch, "speedup_vs_per_op": round(to / tp, 2),
        "bit_equal": bool(np.array_equal(ta.host, tb.host))}
    log(f"bench: rk4 sample persistent {tp:.3f} s per-op {to:.3f} s ({nlaunch} launches)")
    for b in (ta, tb):
        b.free()

    # the full run's final states against the reference's CPU chain of
    # segments: the 32 longest trajectories and 224 more at evenly spaced
    # ranks of the step counts, each at its full length
    if not args.skip_cpu:
        order = np.argsort(steps, kind="stable")[::-1]
        ranks = np.linspace(32, steps.size - 1, 224).astype(np.int64)
        subset = np.concatenate([order[:32], order[ranks]])
        by_len = subset[np.argsort(steps[subset], kind="stable")[::-1]]
        # thread t takes lanes by_len[t::threads] (contiguous in this order)
        subset = np.concatenate([by_len[t::threads] for t in range(threads)])
        cs = steps[subset].astype(np.uint32)
        st3 = state.reshape(steps.size, 96)
        r = ref_program_chain(img, bank, st3[subset].reshape(-1).copy(), cs, 3, threads)
        if r is not None:
            got = full.reshape(steps.size, 96)[subset]
            eq = r[0].reshape(-1, 96) == got
            lanes_bad = int(np.count_nonzero(~eq.all(axis=1)))
            out["lorenz63_fp256_cpu"] = {
                "kind": "reference", "cores": threads, "value": round(int(cs.sum()) / r[3], 1),
                "unit": "steps/s",
                "sample": f"the full run's final states of its 32 longest trajectories and 224 "
                          f"at evenly spaced step ranks, each at its full length ({int(cs.sum())} "
                          f"steps, {int(cs.min())}..{int(cs.max())} a trajectory), the "
                          f"reference's chain of 100-step segments, {threads} threads, "
                          f"{r[3]:.1f} s wall",
                "trajectories_checked": int(subset.size), "lanes_differing": lanes_bad,
                "bit_equal_gpu": lanes_bad == 0 and (r[1] | fl) == fl and r[2] == st}
            log(f"bench: rk4 CPU {int(cs.sum()) / r[3]:.0f} steps/s on {threads} threads; "
                f"{lanes_bad} of {subset.size} full-length trajectories differ")
    for b in (sbuf, bbuf):
        b.free()
    prog.free()

    # Lorenz-96: 8,192 instances, ring 8..256, every ring size in one launch;
    # the planner's knobs tuned on a 1/100 sample of the steps first
    from cftamd import apply_knobs, last_plan, run_multi, tune_multi
    with Device("hip:0", scratch_depth=2048) as d2:
        for fmt, name in ((1, "lorenz96_fp64"), (3, "lorenz96_fp256")):
            groups = odes.lorenz96(8192, seed=args.seed, fmt=fmt)
            progs = [Program(d2, g[1]) for g in groups]
            for p in progs:                 # schedules built outside the timed region
                p.prepare(0x7F)
            tot = sum(int(g[5].sum()) for g in groups)
            comp = sum(int(g[5].sum()) * g[0] for g in groups)
            samp = [np.maximum(1, g[5] // 100).astype(np.uint32) for g in groups]
            knobs, trials = tune_multi(d2, [dict(prog=p, bank=g[2], state=g[4], steps=s2)
                                            for p, g, s2 in zip(progs, groups, samp)])
            parts, kw = apply_knobs([dict(prog=p, bank=g[2], state=g[4].copy(), steps=g[5])
                                     for p, g in zip(progs, groups)], knobs)
            t0 = time.perf_counter()
            fl, _s, widths = run_multi(d2, parts, **kw)
            secs = time.perf_counter() - t0
            plan = last_plan(d2)
            # one instance a ring size - its fewest steps, at full length -
            # against the reference's CPU chain of segments, a thread each
            cpu = None
            if not args.skip_cpu and ref.ok:
                jobs = []
                for g, q in zip(groups, parts):
                    j = int(np.argmin(g[5]))
                    row = g[4].nbytes // g[5].size
                    jobs.append((g, g[4][j * row:(j + 1) * row].copy(), g[5][j:j + 1].copy(),
                                 q["state"][j * row:(j + 1) * row].copy()))
                t0 = time.perf_counter()
                with ThreadPoolExecutor(threads) as ex:
                    rr = list(ex.map(lambda jb: ref_program_chain(jb[0][1], jb[0][2], jb[1],
                                                                  jb[2], fmt, 1), jobs))
                tc = time.perf_counter() - t0
                bad = sum(1 for jb, r in zip(jobs, rr)
                          if r is None or not np.array_equal(r[0], jb[3]) or (r[1] | fl) != fl)
                csteps = sum(int(jb[2][0]) for jb in jobs)
                ccomp = sum(int(jb[2][0]) * jb[0][0] for jb in jobs)
                cpu = {"kind": "reference", "cores": threads, "value": round(csteps / tc, 1),
                       "unit": "steps/s", "component_steps_per_s": round(ccomp / tc, 1),
                       "sample": f"one instance of each of the {len(jobs)} ring sizes (its "
                                 f"fewest steps, {csteps} steps in all) at full length from the "
                                 f"one-launch run, the reference's chain of 20-step segments, "
                                 f"{threads} threads, {tc:.1f} s wall",
                       "instances_checked": len(jobs), "instances_differing": bad,
                       "bit_equal_gpu": bad == 0}
                log(f"bench: {name} CPU {csteps / tc:.0f} steps/s; {bad} of {len(jobs)} differ")
            # on the sample: the same one launch against one batched launch a
            # ring size, bit for bit
            sp, _kw = apply_knobs([dict(prog=p, bank=g[2], state=g[4].copy(), steps=s2)
                                   for p, g, s2 in zip(progs, groups, samp)], knobs)
            t0 = time.perf_counter()
            run_multi(d2, sp, **kw)
            t_multi = time.perf_counter() - t0
            eq = True
            t0 = time.perf_counter()
            outs = []
            for p, g, s2 in zip(progs, groups, samp):
                st2 = g[4].copy()
                p.run_batch(st2, s2, bank=g[2])
                outs.append(st2)
            t_perk = time.perf_counter() - t0
            for o, q in zip(outs, sp):
                eq &= bool(np.array_equal(o, q["state"]))
            stot = int(sum(int(x.sum()) for x in samp))
            # the plain form - one kernel launch an instruction, a batch a
            # ring size - beside the one launch, on steps // 2000 (at least 1)
            samp2 = [np.maximum(1, g[5] // 2000).astype(np.uint32) for g in groups]
            sq, _kw = apply_knobs([dict(prog=p, bank=g[2], state=g[4].copy(), steps=s2)
                                   for p, g, s2 in zip(progs, groups, samp2)], knobs)
            t0 = time.perf_counter()
            run_multi(d2, sq, **kw)
            t_one = time.perf_counter() - t0
            t0 = time.perf_counter()
            nl, eq2 = 0, True
            for p, g, s2, q in zip(progs, groups, samp2, sq):
                st2 = g[4].copy()
                p.run_batch(st2, s2, bank=g[2], mode=L.BATCH_PER_OP)
                nl += int(lib.cftx_batch_launches(d2.handle))
                eq2 &= bool(np.array_equal(st2, q["state"]))
            t_op = time.perf_counter() - t0
            log(f"bench: {name} per-op form {t_op:.2f} s ({nl} launches) vs one launch "
                f"{t_one:.3f} s on steps // 2000")
            for p in progs:
                p.free()
            out[name] = {"instances": 8192, "ring_sizes": len(groups), "total_steps": tot,
                         "seconds": round(secs, 4), "steps_per_s": round(tot / secs, 1),
                         "component_steps_per_s": round(comp / secs, 1), "flags": fl,
                         "launches": 1, "lanes_per_instance_by_ring": {
                             str(g[0]): w for g, w in zip(groups, widths)},
                         "plan": plan, "tuned_knobs": knobs,
                         "tuning_trials_s": [[t, k] for t, k in trials],
                         "sample": {"steps": "steps // 100 (at least 1)", "total_steps": stot,
                                    "one_launch_s": round(t_multi, 4),
                                    "launch_a_ring_s": round(t_perk, 4),
                                    "speedup": round(t_perk / t_multi, 2), "bit_equal": eq},
                         "per_op_sample": {"steps": "steps // 2000 (at least 1)",
                                           "total_steps": int(sum(int(x.sum()) for x in samp2)),
                                           "one_launch_s": round(t_one, 4),
                                           "per_op_s": round(t_op, 4), "per_op_launches": nl,
                                           "speedup_vs_per_op": round(t_op / t_one, 2),
                                           "bit_equal": eq2},
                         "cpu": cpu}
            log(f"bench: {name} {tot / secs / 1e6:.2f} M steps/s ({secs:.2f} s, one launch); "
                f"sample one launch {t_multi:.3f} s vs a launch a ring {t_perk:.3f} s")
    return out


def serve_oracle(pools, reqs, res, samples, threads, ref):
    """The served stream against the reference's CPU library: 32 elements of
    every elementwise job (its first, its last and 30 seeded others), every
    element and the flags of every job of at most 20,000 elements, and lane
    0 of a stratified sample of the ODE jobs (8 Lorenz-63 and 8 Henon-Heiles
    jobs at evenly spaced step ranks, the 6 Lorenz-96 jobs of least work) at
    full length."""
    from cftamd import serve as S
    if not ref.ok:
        return None
    names = {0: "fma", 1: "div", 2: "sqrt"}

    def operands(kind, fmt, pick):
        """a, b, c of a job kind from the pools, at element rows `pick`."""
        esz = 16 if fmt == 2 else 32
        ops = pools.ops[fmt]
        take = lambda k: ops[k].host.reshape(-1, esz)[pick].reshape(-1).copy()
        return (take(3 if kind == 2 else 0), take(1) if kind in (0, 1) else None,
                take(2) if kind == 0 else None)

    t0 = time.perf_counter()
    groups = {}
    for r, _fl, idx, got, _whole in samples:
        groups.setdefault((r["kind"], r["fmt"]), []).append((r, idx, got))
    el_checked = el_bad = 0
    for (kind, fmt), items in sorted(groups.items()):
        esz = 16 if fmt == 2 else 32
        pick = np.concatenate([r["off"] + idx for r, idx, _g in items])
        a, b, c = operands(kind, fmt, pick)
        want, _f, _t = ref.elem(names[kind], fmt, 0, a, b, c, pick.size, threads)
        got = np.concatenate([g.reshape(-1) for _r, _i, g in items])
        diff = (want.reshape(-1, esz) != got.reshape(-1, esz)).any(axis=1)
        el_checked += int(diff.size)
        el_bad += int(diff.sum())

    def whole_job(item):
        r, fl, whole = item
        fmt, n = r["fmt"], r["n"]
        a, b, c = operands(r["kind"], fmt, np.arange(r["off"], r["off"] + n))
        want, wfl, _t = ref.elem(names[r["kind"]], fmt, 0, a, b, c, n, 1)
        return bool(np.array_equal(want, whole)) and wfl == fl

    small = [(r, fl, whole) for r, fl, _i, _g, whole in samples if whole is not None]
    with ThreadPoolExecutor(threads) as ex:
        small_ok = list(ex.map(whole_job, small))

    picks = []
    for kind in (3, 4, 5):
        js = [(r, q) for r, q in zip(reqs, res) if r["kind"] == kind]
        if not js:
            continue
        if kind == 4:
            js.sort(key=lambda rq: rq[0]["k"] * int(np.asarray(rq[0]["steps"])[0]))
            picks += js[:6]
        else:
            js.sort(key=lambda rq: int(np.asarray(rq[0]["steps"])[0]))
            at = sorted(set(np.linspace(0, len(js) - 1, min(8, len(js))).astype(int).tolist()))
            picks += [js[i] for i in at]

    def lane0(rq):
        r, q = rq
        lanes = np.asarray(r["steps"]).size
        row = r["state_bytes"] // lanes
        lo = r["state_off"]
        out = ref_program_chain(pools.image_of(r), pools.bank_of(r),
                                pools.state0[lo:lo + row].copy(),
                                np.asarray(r["steps"], dtype=np.uint32)[:1].copy(), 3, 1)
        return (out is not None and bool(np.array_equal(out[0], pools.states.host[lo:lo + row]))
                and (out[1] | q.flags) == q.flags, int(np.asarray(r["steps"])[0]))

    with ThreadPoolExecutor(threads) as ex:
        lanes = list(ex.map(lane0, picks))
    secs = time.perf_counter() - t0
    return {"kind": "reference", "threads": threads, "seconds": round(secs, 1),
            "elementwise_sampled_elements": el_checked, "elementwise_sampled_differing": el_bad,
            "elementwise_jobs_sampled": len(samples),
            "whole_jobs_checked": len(small), "whole_jobs_differing": small_ok.count(False),
            "whole_job_limit_elements": 20000,
            "ode_lanes_checked": len(lanes), "ode_lanes_differing": sum(not ok for ok, _s in lanes),
            "ode_lane_steps": [s for _ok, s in lanes],
            "bit_exact": el_bad == 0 and all(small_ok) and all(ok for ok, _s in lanes)}


def serve_section(args, threads, ref):
    """The serving loop over a Poisson stream at 2,000 requests/s for one
    second (cftamd/serve.py states the mix and sizes): p50/p99 latency and
    throughput, every job checked against the same job run on its own (every
    elementwise job's output bits in a window of its own), a sample checked
    against the reference's CPU library (serve_oracle), and
    the plain form - one worker, one job a launch, program jobs run whole -
    on the stream's first 20 ms beside the batched loop on the same
    requests."""
    from cftamd import Device
    from cftamd import serve as S
    out = {}
    with Device("hip:0", scratch_depth=2048) as dev:
        reqs = S.requests(seed=args.seed, rate=2000, duration=1.0)
        # every elementwise job writes a window of its own (device memory)
        pools = S.Pools(dev, reqs, seed=args.seed, device_out=True)
        srv = S.Server(dev, workers=4)
        res, wall = S.serve(srv, pools, reqs)
        stats = srv.stats()
        srv.close()
        summ = S.summarize(reqs, res, wall)
        log(f"bench: serve p50 {summ['p50_ms']} ms p99 {summ['p99_ms']} ms over "
            f"{len(reqs)} requests ({wall:.2f} s)")
        samples = []

        def hook(r, q, served):
            esz, n = (16 if r["fmt"] == 2 else 32), r["n"]
            rng = np.random.default_rng(r["off"] * 7919 + n)
            idx = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 30)]))
            samples.append((r, q.flags, idx, served.reshape(n, esz)[idx].copy(),
                            served.copy() if n <= 20000 else None))

        v = S.verify(dev, pools, reqs, res, alone_every=100, hook=hook)
        n_elem = sum(1 for r in reqs if r["kind"] <= 2)
        v["elementwise_jobs"] = n_elem
        log(f"bench: serve verification {v}")
        orc = None if args.skip_cpu else serve_oracle(pools, reqs, res, samples, threads, ref)
        log(f"bench: serve against the reference {orc}")
        del samples
        pre = [r for r in reqs if r["t"] < 0.02]
        pools.states.host[: pools.state0.nbytes] = pools.state0
        pools.states.to_device()
        srv = S.Server(dev, workers=1)
        srv.set_batching(1)
        res_p, wall_p = S.serve(srv, pools, pre)
        srv.close()
        plain = S.summarize(pre, res_p, wall_p)
        same = S.summarize(pre, res[: len(pre)], wall)
        log(f"bench: serve plain form p50 {plain['p50_ms']} ms p99 {plain['p99_ms']} ms on "
            f"{len(pre)} requests; batched on the same: p50 {same['p50_ms']} p99 "
            f"{same['p99_ms']}")
        pools.free()
        out = {"workload": "Poisson 2,000 requests/s for 1 s, seed %d" % args.seed,
               "requests": summ["requests"], "p50_ms": summ["p50_ms"],
               "p99_ms": summ["p99_ms"], "wall_s": summ["wall_s"],
               "served_req_s": summ["throughput_req_s"], "by_kind": summ["by_kind"],
               "elements": summ["elements"], "ode_steps": summ["ode_steps"],
               "statuses_ok": summ["statuses_ok"], "server": stats, "verification": v,
               "vs_reference": orc,
               "plain_form_prefix": {"requests": len(pre), "p50_ms": plain["p50_ms"],
                                     "p99_ms": plain["p99_ms"], "wall_s": plain["wall_s"],
                                     "batched_p50_ms": same["p50_ms"],
                                     "batched_p99_ms": same["p99_ms"]}}
    return out


def cpu_baseline(a, b, c, gpu_out, n, threads, ref, info):
    """The reference's CPU path (its own C library, compiled from its sources
    into oracle/_ref by oracle/Makefile) on the same operands: one software
    device per thread over contiguous slices, every mode, every result
    compared; then one thread on the first 2^18 elements (RNE) for the
    single-core rate.  Falls back to the oracle restatement when the
    reference library is not present."""
    port_path = os.path.join(HERE, "oracle", "_build", "libcftoracle.so")
    V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)
    if ref.ok:
        kind = "reference"

        def run_all(rnd, nn, th):
            return ref.elem("fma", 3, rnd, a, b, c, nn, th)[0]
    elif os.path.exists(port_path):
        kind = "port"
        lib = ctypes.CDLL(port_path)
        lib.oc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                               ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P]

        def run_all(rnd, nn, th):
            out = np.empty(nn * 32, dtype=np.uint8)
            f = ctypes.c_uint32()
            lib.oc_run(3, 0, rnd, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                       out.ctypes.data, nn, 0, th, ctypes.byref(f))
            return out
    else:
        return None, None
    mismatches = 0
    t0 = time.perf_counter()
    for rnd in range(5):
        out = run_all(rnd, n, threads)
        mismatches += int(np.count_nonzero(
            (out.reshape(-1, 32) != gpu_out[rnd][: n * 32].reshape(-1, 32)).any(axis=1)))
    secs = time.perf_counter() - t0
    n1 = min(n, 1 << 18)
    t1 = time.perf_counter()
    one = run_all(0, n1, 1)
    s1 = time.perf_counter() - t1
    mismatches += int(np.count_nonzero(
        (one.reshape(-1, 32) != gpu_out[0][: n1 * 32].reshape(-1, 32)).any(axis=1)))
    return {"value": round(5 * n / secs / 1e9, 6), "unit": "Gelem/s", "cores": threads,
            "kind": kind,
            "sample": f"binary256 FMA, the first {n} elements of the benchmark's operands, "
                      f"all five rounding modes ({5 * n} elements), {threads} threads, one "
                      f"software device per thread, {secs:.1f} s wall; single core: the "
                      f"first {n1} elements, RNE, {s1:.2f} s",
            "single_core": {"value": round(n1 / s1 / 1e9, 6), "unit": "Gelem/s",
                            "ns_per_elem": round(s1 / n1 * 1e9, 1)},
            "all_cores_ns_per_elem": round(secs / (5 * n) * 1e9 * threads, 1),
            "nproc": info["nproc"], "cpus_available": info["cpus_available"],
            "physical_cores": info["physical_cores"], "cpu_model": info["cpu_model"]}, \
        mismatches


def check_op(lib, dev, bufs, n, name, rnd, ref, threads, sample):
    """One binary256 op over the resident operands against the reference's
    CPU library: every element (add, mul) or every 16th (div, sqrt), and for
    the sample the GPU's own call over the gathered elements beside it, so
    flags are compared exactly.  Returns (elements checked, differing)."""
    fl = ctypes.c_uint32()
    a, b, c, sa = bufs[0], bufs[1], bufs[2], bufs[4]
    if name == "add":
        st = lib.cft_run(dev.handle, 1, 3, rnd, a.data_ptr, None, c.data_ptr, bufs[3].data_ptr,
                         n, ctypes.byref(fl), None)
    elif name == "mul":
        st = lib.cft_run(dev.handle, 3, 3, rnd, a.data_ptr, b.data_ptr, None, bufs[3].data_ptr,
                         n, ctypes.byref(fl), None)
    elif name == "div":
        st = lib.cft_div(dev.handle, 3, rnd, a.data_ptr, b.data_ptr, bufs[3].data_ptr, n,
                         ctypes.byref(fl), None)
    else:
        st = lib.cft_sqrt(dev.handle, 3, rnd, sa.data_ptr, bufs[3].data_ptr, n,
                          ctypes.byref(fl), None)
    assert st == 0, lib.cft_last_error()
    bufs[3].from_device()
    got = bufs[3].host.reshape(n, 32)
    if not sample:
        want, wfl, _t = ref.elem(name, 3, rnd, a.host, b.host, c.host, n, threads)
        bad = int(np.count_nonzero((want.reshape(n, 32) != got).any(axis=1)))
        return n, bad + (0 if wfl == fl.value else 1)
    pick = np.arange(0, n, 16)
    take = lambda x: x.host.reshape(n, 32)[pick].reshape(-1).copy()
    xa, xb = (take(sa) if name == "sqrt" else take(a)), take(b)
    want, wfl, _t = ref.elem(name, 3, rnd, xa, xb, None, pick.size, threads)
    bad = int(np.count_nonzero((want.reshape(-1, 32) != got[pick]).any(axis=1)))
    if name == "div":
        g2, gfl = dev.div(3, rnd, xa, xb)
    else:
        g2, gfl = dev.sqrt(3, rnd, xa)
    bad += int(np.count_nonzero((want.reshape(-1, 32) != g2.reshape(-1, 32)).any(axis=1)))
    return pick.size, bad + (0 if wfl == gfl else 1)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(elements and lanes shard by index with no exchange but a flag OR)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None

    from cftamd import Device, workload, _lib
    lib = _lib.lib()
    info = host_info()
    threads = cpu_threads(args, info)
    ref = RefCPU()
    log(f"bench: host {info}, {threads} CPU threads, reference library "
        f"{'present' if ref.ok else 'absent'}")
    n = args.n
    log(f"bench: generating {n} binary256 operand triples (seed {args.seed})")
    t0 = time.perf_counter()
    dev = Device("hip:0")
    bufs = [dev.alloc(n * 32) for _ in range(5)]
    workload.elementwise(2, n, seed=args.seed, threads=16,
                         out=(bufs[0].host, bufs[1].host, bufs[2].host))
    # sqrt reads |a|
    bufs[4].host[:] = bufs[0].host
    bufs[4].host.reshape(n, 32)[:, 31] &= 0x7F
    for bf in bufs[:3] + bufs[4:]:
        bf.to_device()
    log(f"bench: operands resident in HBM ({time.perf_counter() - t0:.1f} s)")

    fl = ctypes.c_uint32()

    def fma_step():
        for rnd in range(5):
            st = lib.cft_run(dev.handle, 0, 3, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                             bufs[2].data_ptr, bufs[3].data_ptr, n, ctypes.byref(fl), None)
            if st != 0:
                raise RuntimeError(lib.cft_last_error().decode())

    for _ in range(args.warmup):
        fma_step()
    _lib.check(lib.cftx_kernel_timer(dev.handle, 1), "cftx_kernel_timer")
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fma_step()
    sync()
    secs = time.perf_counter() - t0
    kms = ctypes.c_double()
    klaunch = ctypes.c_uint64()
    lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(klaunch))
    lib.cftx_kernel_timer(dev.handle, 0)
    ms_step = secs * 1e3 / args.steps
    value = 5 * n * args.steps / secs / 1e9
    kavg_s = kms.value / 1e3 / max(1, klaunch.value)
    achieved = FMA_BYTES * n / kavg_s / 1e9
    log(f"bench: FMA {value:.3f} Gelem/s call-to-call, kernel {kavg_s * 1e3:.3f} ms "
        f"per launch ({n / kavg_s / 1e9:.3f} Gelem/s in-kernel)")
    rates = issue_rates()
    log(f"bench: issue rates {rates}")

    # the outputs of every mode, for the CPU comparison
    gpu_out = []
    for rnd in range(5):
        st = lib.cft_run(dev.handle, 0, 3, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                         bufs[2].data_ptr, bufs[3].data_ptr, n, ctypes.byref(fl), None)
        assert st == 0
        bufs[3].from_device()
        gpu_out.append(bufs[3].host.copy())

    ops = {}
    if not args.skip_ops:
        for name in ("add", "mul", "div", "sqrt"):
            def op_call(rnd, name=name):
                if name == "add":
                    return lib.cft_run(dev.handle, 1, 3, rnd, bufs[0].data_ptr, None,
                                       bufs[2].data_ptr, bufs[3].data_ptr, n, None, None)
                if name == "mul":
                    return lib.cft_run(dev.handle, 3, 3, rnd, bufs[0].data_ptr,
                                       bufs[1].data_ptr, None, bufs[3].data_ptr, n, None, None)
                if name == "div":
                    return lib.cft_div(dev.handle, 3, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                                       bufs[3].data_ptr, n, None, None)
                return lib.cft_sqrt(dev.handle, 3, rnd, bufs[4].data_ptr, bufs[3].data_ptr,
                                    n, None, None)
            # one untimed call first: a kernel's first launch also loads its
            # code object, which would land between the timer's events
            assert op_call(0) == 0, lib.cft_last_error()
            lib.cftx_kernel_timer(dev.handle, 1)
            t1 = time.perf_counter()
            for rnd in range(5):
                st = op_call(rnd)
                assert st == 0, lib.cft_last_error()
            dt = time.perf_counter() - t1
            lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(klaunch))
            lib.cftx_kernel_timer(dev.handle, 0)
            ks = kms.value / 1e3 / max(1, klaunch.value)
            ops[name] = {"Gelem_s": round(5 * n / dt / 1e9, 4),
                         "kernel_ms": round(ks * 1e3, 4),
                         "kernel_GBs": round(BYTES[name] * n / ks / 1e9, 1)}
            if name == "sqrt":
                ops[name]["operand"] = "|a| (the sign bit of a cleared)"
            if not args.skip_cpu and ref.ok:
                sample = name in ("div", "sqrt")
                t1 = time.perf_counter()
                tot = bad = 0
                for rnd in range(5):
                    k, b_ = check_op(lib, dev, bufs, n, name, rnd, ref, threads, sample)
                    tot += k
                    bad += b_
                ops[name]["vs_reference_cpu"] = {
                    "elements_checked": tot, "differing": bad, "bit_exact": bad == 0,
                    "checked": ("every 16th element of each of the 5 modes, with the GPU's "
                                "own call over the gathered elements for the flags")
                    if sample else "every element of all 5 modes, flags included",
                    "seconds": round(time.perf_counter() - t1, 1)}
            log(f"bench: {name} {ops[name]}")
        # reductions over the same resident binary256 operands: the tree over
        # the whole array, and the same over segments (one result a segment)
        red = {}
        for name, op, seg, b_ptr in (("sum", 24, None, None),
                                     ("dot", 25, None, bufs[1].data_ptr),
                                     ("maxall", 31, None, None),
                                     ("sum_seg4096", 24, 4096, None),
                                     ("sum_seg16", 24, 16, None)):
            nres = 1 if seg is None else n // seg
            rbuf = dev.alloc(nres * 32)
            f2 = ctypes.c_uint32()

            def red_call():
                if seg is None:
                    st = lib.cft_reduce(dev.handle, op, 3, 0, bufs[0].data_ptr, b_ptr,
                                        rbuf.data_ptr, n, ctypes.byref(f2), None)
                else:
                    st = lib.cft_reduce_seg(dev.handle, op, 3, 0, bufs[0].data_ptr, b_ptr,
                                            rbuf.data_ptr, n, seg, ctypes.byref(f2), None)
                assert st == 0, lib.cft_last_error()

            red_call()
            lib.cftx_kernel_timer(dev.handle, 1)
            t1 = time.perf_counter()
            for _ in range(5):
                red_call()
            dt = (time.perf_counter() - t1) / 5
            lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(klaunch))
            lib.cftx_kernel_timer(dev.handle, 0)
            ks = kms.value / 1e3 / max(1, klaunch.value)
            rbytes = (64 if op == 25 else 32) * n
            red[name] = {"Gelem_s": round(n / dt / 1e9, 3), "kernel_ms": round(ks * 1e3, 4),
                         "kernel_Gelem_s": round(n / ks / 1e9, 3),
                         "kernel_GBs": round(rbytes / ks / 1e9, 1), "results": nres,
                         "flags": f2.value}
            log(f"bench: reduce {name} {red[name]}")
            rbuf.free()
        ops["reduce_fp256"] = red
        # the binary128 configuration: FMA RNE over 2^20 elements (seed 0),
        # checked against the reference CPU path on all of them
        n1 = 1 << 20
        b1 = [dev.alloc(n1 * 16) for _ in range(4)]
        workload.elementwise(1, n1, seed=args.seed, threads=16,
                             out=(b1[0].host, b1[1].host, b1[2].host))
        for bf in b1[:3]:
            bf.to_device()
        f1 = ctypes.c_uint32()
        lib.cft_run(dev.handle, 0, 2, 0, b1[0].data_ptr, b1[1].data_ptr, b1[2].data_ptr,
                    b1[3].data_ptr, n1, ctypes.byref(f1), None)
        # a call here is ~30 us, so the call-to-call rate is timed without
        # the kernel timer (its two events and their read-back are ~8 us of
        # it) and the kernel's duration in a second pass with it
        t1 = time.perf_counter()
        for _ in range(20):
            lib.cft_run(dev.handle, 0, 2, 0, b1[0].data_ptr, b1[1].data_ptr, b1[2].data_ptr,
                        b1[3].data_ptr, n1, ctypes.byref(f1), None)
        dt = time.perf_counter() - t1
        lib.cftx_kernel_timer(dev.handle, 1)
        for _ in range(20):
            lib.cft_run(dev.handle, 0, 2, 0, b1[0].data_ptr, b1[1].data_ptr, b1[2].data_ptr,
                        b1[3].data_ptr, n1, ctypes.byref(f1), None)
        lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(klaunch))
        lib.cftx_kernel_timer(dev.handle, 0)
        ks = kms.value / 1e3 / max(1, klaunch.value)
        b1[3].from_device()
        ops["fma_fp128_c1"] = {"n": n1, "Gelem_s": round(20 * n1 / dt / 1e9, 4),
                               "kernel_ms": round(ks * 1e3, 4),
                               "kernel_Gelem_s": round(n1 / ks / 1e9, 4),
                               "kernel_GBs": round(64 * n1 / ks / 1e9, 1), "flags": f1.value}
        if not args.skip_cpu and ref.ok:
            r = ref.elem("fma", 2, 0, b1[0].host, b1[1].host, b1[2].host, n1, threads)
            ops["fma_fp128_c1"]["bit_exact_vs_cpu"] = bool(
                np.array_equal(r[0], b1[3].host) and r[1] == f1.value)
            ops["fma_fp128_c1"]["cpu_Gelem_s"] = round(n1 / r[2] / 1e9, 5)
            ops["fma_fp128_c1"]["cpu_threads"] = threads
        log(f"bench: fp128 FMA {ops['fma_fp128_c1']}")
        for bf in b1:
            bf.free()

    rk4 = None
    if not args.skip_rk4:
        rk4 = rk4_section(dev, lib, args, threads, ref)

    serve = None
    if not args.skip_serve:
        serve = serve_section(args, threads, ref)

    cpu = None
    bit_exact = None
    if not args.skip_cpu:
        a = bufs[0].host.copy()
        b = bufs[1].host.copy()
        c = bufs[2].host.copy()
        log(f"bench: CPU baseline on {threads} threads")
        cpu, mism = cpu_baseline(a, b, c, gpu_out, n, threads, ref, info)
        if cpu is not None:
            bit_exact = mism == 0
            log(f"bench: CPU {cpu['value']} Gelem/s ({cpu['kind']}), mismatches {mism}")

    stream = stream_probe(n)
    if stream:
        log(f"bench: torch.addcmul on the same bytes streams {stream:.0f} GB/s; the FMA kernel "
            f"moves {achieved:.0f} GB/s ({achieved / stream:.2f} of it)")
    kernel_gelem = n / kavg_s / 1e9
    valu = None
    if rates:
        wave_insns = FMA_VALU_PER_ELEM * (n / 64) / kavg_s
        mad_rate = rates["wave_insns_per_s"]["v_mad_u64_u32"] * 1e9
        valu = {
            "valu_issue": round(wave_insns / VALU_PEAK_WAVE_INSNS, 4),
            "valu_issue_vs_4cycle_rate": round(wave_insns / VALU_4CYCLE_WAVE_INSNS, 4),
            "valu_issue_vs_measured_add": round(
                wave_insns / (rates["wave_insns_per_s"]["v_add_u32"] * 1e9), 4),
            "valu_wave_insns_s": round(wave_insns, 1),
            "valu_instr_per_elem": FMA_VALU_PER_ELEM,
            "valu_cycles_per_instr_per_simd": FMA_CYCLES_PER_VALU,
            "alg_valu_frac": round(kernel_gelem * 1e9 * FMA_MACS / 64 / mad_rate, 4),
            "peak_wave_insns_s": VALU_PEAK_WAVE_INSNS,
            "mad_u64_u32_wave_insns_s": round(mad_rate, 1),
            "issue_rates": rates,
            "valu_source": "valu_instr_per_elem and cycles a VALU instruction: rocprofv3 "
                           "SQ_INSTS_VALU, SQ_WAVES, GRBM_GUI_ACTIVE of this kernel on this "
                           "workload (profiles/valu_issue_rates.md); valu_issue = that x "
                           "elements/64 / kernel time / (256 CUs x 4 SIMDs x 2.4 GHz / 2 "
                           "cycles); alg_valu_frac = the exact product's 64 MACs an element "
                           "at the measured v_mad_u64_u32 issue rate: (elem/s x 64 / 64 "
                           "lanes) / that rate"}
        log(f"bench: VALU {valu['valu_issue']} of the 2-cycle peak "
            f"({valu['valu_issue_vs_4cycle_rate']} of the 4-cycle rate), "
            f"alg {valu['alg_valu_frac']}")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(FMA_TRAFFIC_PER_ELEM * n),
            "kernel_ms": round(kavg_s * 1e3, 4),
            "kernel_Gelem_s": round(kernel_gelem, 4),
            "bytes_per_elem": FMA_BYTES,
            "stream_GBs": round(stream, 1) if stream else None,
            "frac_of_stream": round(achieved / stream, 4) if stream else None}
    if valu:
        roof.update(valu)
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "binary256 FMA over the elementwise configuration's "
                               f"operands, N={n}, seed {args.seed}, 5 rounding modes a step",
                   "elements_per_step": 5 * n, "format": "binary256", "resident": True},
        "roofline": roof,
        "cpu_baseline": cpu,
        "bit_exact_vs_cpu": bit_exact,
        "ops": ops,
        "rk4": rk4,
        "serve": serve,
    }
    for bf in bufs:
        bf.free()
    dev.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
