"""Run one of the benchmark's workloads once, for profilers and quick checks.

    python -m cftamd.run_workload fma   [--n N] [--seed S]   binary256 FMA, 5 modes
    python -m cftamd.run_workload l63   [--seed S]           Lorenz-63 ensemble, one launch
    python -m cftamd.run_workload l96   [--fmt 1|3] [--seed S]  Lorenz-96, one launch
    python -m cftamd.run_workload serve [--seconds T] [--seed S]  the serving stream
    python -m cftamd.run_workload reduce [--n N] [--fmt 2|3] [--seg S] [--op OP] [--seed S]
                                        a reduction (SUM 24, DOT 25, SUMSQ 28,
                                        SUMABS 29, MAXALL 31) over the FMA
                                        operands (binary128 or binary256), 5 modes

Each builds its inputs as bench.py does (same generators and seeds), runs
the workload once on GPU 0 with every input resident in device memory, and
prints one line of timing.  Nothing is checked here; the tests and bench.py
check the bits.
"""

import argparse
import ctypes
import sys
import time

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


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cftamd.run_workload")
    ap.add_argument("workload", choices=["fma", "l63", "l96", "serve", "reduce"])
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--fmt", type=int, default=3, choices=[1, 2, 3])
    ap.add_argument("--seg", type=int, default=0)
    ap.add_argument("--op", type=int, default=24, choices=[24, 25, 28, 29, 31])
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    {"fma": fma, "l63": l63, "l96": l96, "serve": serve,
     "reduce": reduce}[args.workload](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
