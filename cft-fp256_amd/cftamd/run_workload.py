"""Run one of the benchmark's workloads once, for profilers and quick checks.

    python -m cftamd.run_workload fma   [--n N] [--seed S]   binary256 FMA, 5 modes
    python -m cftamd.run_workload divsqrt [--n N] [--seed S] binary256 division, then
                                        square root of |a|, 5 modes each
    python -m cftamd.run_workload l63   [--seed S]           Lorenz-63 ensemble, one launch
    python -m cftamd.run_workload l96   [--fmt 1|3] [--seed S]  Lorenz-96, one launch
    python -m cftamd.run_workload serve [--seconds T] [--seed S]  the serving stream
    python -m cftamd.run_workload reduce [--n N] [--fmt 2|3] [--seg S] [--op OP] [--seed S]
                                        a reduction (SUM 24, DOT 25, SUMSQ 28,
                                        SUMABS 29, MAXALL 31) over the FMA
                                        operands (binary128 or binary256), 5 modes
    python -m cftamd.run_workload issue    the issue-rate probes (lib/libcftprobe.so)
    python -m cftamd.run_workload latency  one wavefront alone: its issue interval and
                                        dependent latency (lib/libcftprobe.so), a
                                        binary256 FMA / ADD / steered MUL chained in
                                        registers, the wide runner's chains of 16
                                        dependent FMA, MUL, ADD and COPYSIGN a trip
                                        (two lanes an instance), and one Lorenz-63
                                        trajectory of 20,000 steps
    python -m cftamd.run_workload kernels  the other kernels once each: k_misc
                                        (binary256 MIN over the FMA operands,
                                        --n elements), k_seq (one Lorenz-63
                                        segment of 100 steps over 65,536 lanes),
                                        k_seq_batch (the ensemble's first 16,384
                                        trajectories at steps // 100, one lane an
                                        instance) and k_seq_op (the plain form
                                        on 4,096 of them at steps // 1,000)

Each builds its inputs as bench.py does (same generators and seeds), runs
the workload once on GPU 0 with every input resident in device memory, and
prints one line of timing.  Nothing is checked here; the tests and bench.py
check the bits.
"""

import argparse
import ctypes
import sys
import time

import numpy as np

from . import Device, Program, _lib, odes, run_multi, workload


def fma(args):
    """binary256 (--fmt 3): the five rounding modes, one call each, over the
    elementwise configuration's operands.  binary128 (--fmt 2): twenty RNE
    calls over the binary128 configuration's operands (one untimed first)."""
    lib = _lib.lib()
    n = args.n
    fmt = 2 if args.fmt == 2 else 3
    esz = 16 if fmt == 2 else 32
    with Device("hip:0") as dev:
        bufs = [dev.alloc(n * esz) for _ in range(4)]
        workload.elementwise(fmt - 1, n, seed=args.seed, threads=16,
                             out=(bufs[0].host, bufs[1].host, bufs[2].host))
        for b in bufs[:3]:
            b.to_device()
        fl = ctypes.c_uint32()
        modes = [0] * 20 if fmt == 2 else list(range(5))

        def call(rnd):
            _lib.check(lib.cft_run(dev.handle, 0, fmt, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                                   bufs[2].data_ptr, bufs[3].data_ptr, n, ctypes.byref(fl),
                                   None), "cft_run")
        if fmt == 2:
            call(0)
        t0 = time.perf_counter()
        for rnd in modes:
            call(rnd)
        dt = time.perf_counter() - t0
        k = len(modes) * n
        print(f"fma: {k} elements in {dt:.4f} s ({k / dt / 1e9:.3f} Gelem/s), flags {fl.value}")
        for b in bufs:
            b.free()


def divsqrt(args):
    """binary256 division then square root (of |a|), the five rounding modes
    each, over the elementwise configuration's operands."""
    lib = _lib.lib()
    n = args.n
    with Device("hip:0") as dev:
        bufs = [dev.alloc(n * 32) for _ in range(4)]
        workload.elementwise(2, n, seed=args.seed, threads=16,
                             out=(bufs[0].host, bufs[1].host, bufs[2].host))
        bufs[2].host[:] = bufs[0].host                   # sqrt reads |a|
        bufs[2].host.reshape(n, 32)[:, 31] &= 0x7F
        for b in bufs[:3]:
            b.to_device()
        fl = ctypes.c_uint32()
        for name in ("div", "sqrt"):
            t0 = time.perf_counter()
            for rnd in range(5):
                if name == "div":
                    st = lib.cft_div(dev.handle, 3, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                                     bufs[3].data_ptr, n, ctypes.byref(fl), None)
                else:
                    st = lib.cft_sqrt(dev.handle, 3, rnd, bufs[2].data_ptr, bufs[3].data_ptr, n,
                                      ctypes.byref(fl), None)
                _lib.check(st, "cft_" + name)
            dt = time.perf_counter() - t0
            print(f"{name}: {5 * n} elements in {dt:.4f} s ({5 * n / dt / 1e9:.3f} Gelem/s), "
                  f"flags {fl.value}")
        for b in bufs:
            b.free()


def reduce(args):
    lib = _lib.lib()
    n, fmt, esz = args.n, args.fmt, _lib.FORMAT_BYTES[args.fmt]
    with Device("hip:0") as dev:
        a = dev.alloc(n * esz)
        scratch = [dev.alloc(n * esz) for _ in range(2)]
        workload.elementwise(1 if fmt == 2 else 2, n, seed=args.seed, threads=16,
                             out=(a.host, scratch[0].host, scratch[1].host))
        a.to_device()
        scratch[0].to_device()
        bptr = scratch[0].data_ptr if args.op == 25 else None
        nres = n // args.seg if args.seg else 1
        d = dev.alloc(nres * esz)
        fl = ctypes.c_uint32()
        t0 = time.perf_counter()
        for rnd in range(5):
            if args.seg:
                st = lib.cft_reduce_seg(dev.handle, args.op, fmt, rnd, a.data_ptr, bptr,
                                        d.data_ptr, n, args.seg, ctypes.byref(fl), None)
            else:
                st = lib.cft_reduce(dev.handle, args.op, fmt, rnd, a.data_ptr, bptr, d.data_ptr,
                                    n, ctypes.byref(fl), None)
            _lib.check(st, "cft_reduce")
        dt = time.perf_counter() - t0
        print(f"reduce op {args.op} fmt {fmt} seg {args.seg}: {5 * n} elements in {dt:.4f} s "
              f"({5 * n / dt / 1e9:.3f} Gelem/s), flags {fl.value}")
        for b in [a, d] + scratch:
            b.free()


def l63(args):
    img, bank, state, steps = odes.lorenz63(65536, seed=args.seed, fmt=3)
    with Device("hip:0") as dev:
        prog = Program(dev, img)
        sbuf = dev.alloc(state.nbytes)
        sbuf.host[:] = state
        sbuf.to_device()
        t0 = time.perf_counter()
        fl, st = prog.run_batch(sbuf, steps, bank=bank)
        dt = time.perf_counter() - t0
        print(f"l63: {int(steps.sum())} steps in {dt:.3f} s "
              f"({steps.sum() / dt / 1e6:.1f} M steps/s), flags {fl}, "
              f"lanes an instance {_lib.lib().cftx_batch_width(dev.handle)}")
        sbuf.free()
        prog.free()


def l96(args):
    with Device("hip:0", scratch_depth=2048) as dev:
        groups = odes.lorenz96(8192, seed=args.seed, fmt=args.fmt)
        progs = [Program(dev, g[1]) for g in groups]
        for p in progs:
            p.prepare(0x7F)
        parts = [dict(prog=p, bank=g[2], state=g[4].copy(), steps=g[5])
                 for p, g in zip(progs, groups)]
        tot = sum(int(g[5].sum()) for g in groups)
        t0 = time.perf_counter()
        fl, _s, widths = run_multi(dev, parts)
        dt = time.perf_counter() - t0
        print(f"l96 fmt {args.fmt}: {tot} steps in {dt:.3f} s ({tot / dt / 1e6:.2f} M steps/s), "
              f"flags {fl}, lanes an instance by ring {widths}")
        for p in progs:
            p.free()


def serve(args):
    from . import serve as S
    with Device("hip:0", scratch_depth=2048) as dev:
        reqs = S.requests(seed=args.seed, rate=2000, duration=args.seconds)
        pools = S.Pools(dev, reqs, seed=args.seed)
        srv = S.Server(dev, workers=4)
        res, wall = S.serve(srv, pools, reqs)
        srv.close()
        summ = S.summarize(reqs, res, wall)
        print(f"serve: {len(reqs)} requests in {wall:.2f} s, p50 {summ['p50_ms']} ms, "
              f"p99 {summ['p99_ms']} ms")
        pools.free()


def issue(args):
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib",
                        "libcftprobe.so")
    P = ctypes.CDLL(path)
    P.cftp_issue_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
    for kind, name in enumerate(("v_add_u32", "v_mad_u64_u32", "v_mul_lo_u32",
                                 "v_mul_hi_u32", "v_add_co_u32+v_addc_co_u32")):
        ms, wi = ctypes.c_double(), ctypes.c_double()
        if P.cftp_issue_rate(kind, 2048, 16384, 5, ctypes.byref(ms), ctypes.byref(wi)) != 0:
            raise RuntimeError("the issue-rate probe failed")
        print(f"issue {name}: {wi.value / (ms.value * 1e-3) / 1e9:.1f} G wave instructions/s "
              f"({ms.value:.3f} ms a launch)")


def _probe():
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib",
                        "libcftprobe.so")
    return ctypes.CDLL(path)


def chain_program(op, per_trip=16, fmt=3):
    """x <- op(x, ...) per_trip times a trip of the segment loop, one value in
    flight: FMA x*K0 + K1, MUL x*K0, ADD x + K1, COPYSIGN(x, K0); K0 =
    RN(1.0000001), K1 = RN(1e-30)."""
    from . import programs as P
    one = {"fma": P.alu(P.FMA, 1, 1, P.K(0), P.K(1)), "mul": P.alu(P.MUL, 1, 1, P.K(0)),
           "add": P.alu(P.ADD, 1, 1, 0, P.K(1)), "copysign": P.alu(6, 1, 1, P.K(0))}[op]
    ins = [P.ldl(1, 0), P.repeat(1)] + [one] * per_trip + [P.endrep(), P.stl(1, 0), P.halt()]
    return P.image(fmt, ins, consts=[P.dec(fmt, "1.0000001"), P.dec(fmt, "1e-30")],
                   scratch_in=1, scratch_out=1)


def latency(args):
    from . import programs as P
    pr = _probe()
    pr.cftp_lone_issue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
    pr.cftp_chain_latency.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    for kind, name in ((0, "v_add_u32"), (1, "v_mad_u64_u32"), (2, "ds_bpermute_b32"),
                       (3, "v_mov_b32 row_shr:1 (DPP)")):
        for ch in (1, 2, 4, 8):
            c = ctypes.c_double()
            if pr.cftp_lone_issue(kind, ch, 20000, ctypes.byref(c)) != 0:
                raise RuntimeError("the lone-issue probe failed")
            print(f"lone wave {name}, {ch} chain(s): {c.value:.2f} clocks an instruction")
    words = np.frombuffer(b"".join(int(v).to_bytes(32, "little") for v in
                                   (P.dec(3, "1.5"), P.dec(3, "1.0000001"), P.dec(3, "1e-30"))),
                          np.uint32).copy()
    out = np.zeros(9, np.uint32)
    for kind, name in ((0, "FMA"), (1, "ADD (two-term path)"), (2, "MUL (steered to FMA)")):
        us = ctypes.c_double()
        if pr.cftp_chain_latency(kind, 20000, words.ctypes.data, out.ctypes.data,
                                 ctypes.byref(us)) != 0:
            raise RuntimeError("the chain probe failed")
        print(f"register chain {name}: {us.value:.3f} us an operation")
    pr.cftp_lp_chain_latency.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_double)]
    out10 = np.zeros(10, np.uint32)
    us = ctypes.c_double()
    if pr.cftp_lp_chain_latency(20000, words.ctypes.data, out10.ctypes.data,
                                ctypes.byref(us)) != 0:
        raise RuntimeError("the eight-lane chain probe failed")
    print(f"register chain FMA over eight lanes (fx_lp.hpp): {us.value:.3f} us an operation "
          f"({int(out10[9])} handed back)")
    with Device("hip:0") as dev:
        for op in ("fma", "mul", "add", "copysign"):
            prog = Program(dev, chain_program(op))
            st = odes.widen(np.array([1.5]), 3).copy()
            prog.run_batch(st.copy(), np.array([100], np.uint32), lanes_per_instance=2)
            t0 = time.perf_counter()
            prog.run_batch(st, np.array([10000], np.uint32), lanes_per_instance=2)
            dt = time.perf_counter() - t0
            print(f"wide runner chain {op}: {dt / 10000 / 16 * 1e6:.3f} us an operation "
                  f"(16 a trip, 2 lanes an instance)")
            prog.free()
        img, bank, state, steps = odes.lorenz63(1, seed=args.seed, fmt=3)
        prog = Program(dev, img)
        for w in (4, 8):
            s = state.copy()
            prog.run_batch(s, np.array([200], np.uint32), bank=bank, lanes_per_instance=w)
            s = state.copy()
            t0 = time.perf_counter()
            prog.run_batch(s, np.array([20000], np.uint32), bank=bank, lanes_per_instance=w)
            dt = time.perf_counter() - t0
            print(f"lone Lorenz-63, {w} lanes an instance: {dt / 20000 * 1e6:.2f} us a step")
        prog.free()


def kernels(args):
    lib = _lib.lib()
    n = args.n
    with Device("hip:0") as dev:
        bufs = [dev.alloc(n * 32) for _ in range(3)]
        workload.elementwise(2, n, seed=args.seed, threads=16,
                             out=(bufs[0].host, bufs[1].host, bufs[2].host))
        for b in bufs[:2]:
            b.to_device()
        fl = ctypes.c_uint32()
        for _ in range(3):
            st = lib.cft_run(dev.handle, 7, 3, 0, bufs[0].data_ptr, bufs[1].data_ptr, None,
                             bufs[2].data_ptr, n, ctypes.byref(fl), None)
            assert st == 0, lib.cft_last_error()
        print(f"k_misc: binary256 MIN over {n} elements x 3, flags {fl.value}")
        for b in bufs:
            b.free()
        img, bank, state, steps = odes.lorenz63(65536, seed=args.seed, fmt=3)
        prog = Program(dev, img)
        lanes = steps.size
        za = np.zeros(lanes * 32, np.uint8)
        sout = np.empty_like(state)
        f, s = prog.run(lanes, za, bank=bank, scratch_in=state, scratch_out=sout)
        print(f"k_seq: one 100-step segment over {lanes} lanes, flags {f}")
        m = 16384
        st2 = state[: m * 96].copy()
        f, s = prog.run_batch(st2, np.maximum(1, steps[:m] // 100), bank=bank,
                              lanes_per_instance=1)
        print(f"k_seq_batch: {m} trajectories at steps // 100, one lane an instance, flags {f}")
        m = 4096
        st3 = state[: m * 96].copy()
        f, s = prog.run_batch(st3, np.maximum(1, steps[:m] // 1000), bank=bank,
                              mode=_lib.BATCH_PER_OP)
        print(f"k_seq_op: {m} trajectories at steps // 1000, "
              f"{int(lib.cftx_batch_launches(dev.handle))} launches, flags {f}")
        prog.free()


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cftamd.run_workload")
    ap.add_argument("workload", choices=["fma", "divsqrt", "l63", "l96", "serve", "reduce",
                                         "issue",
                                         "latency", "kernels"])
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--fmt", type=int, default=3, choices=[1, 2, 3])
    ap.add_argument("--seg", type=int, default=0)
    ap.add_argument("--op", type=int, default=24, choices=[24, 25, 28, 29, 31])
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    {"fma": fma, "divsqrt": divsqrt, "l63": l63, "l96": l96, "serve": serve, "reduce": reduce,
     "issue": issue,
     "latency": latency, "kernels": kernels}[args.workload](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
