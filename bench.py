#!/usr/bin/env python3
"""Benchmark of the MI355X engine on the reference's headline workload.

Workload (one GPU): binary256 operands of the elementwise configuration -
N = 2^24 elements from seed 0, exponents over the full range, 3% subnormals,
2% specials, 5% FMA cases constructed near a rounding boundary.  One step is
the fused multiply-add over all N elements in each of the five rounding
modes (five cft_run calls), with the operands resident in HBM (cft_alloc
buffers) before the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line: the FMA rate (Gelem/s), the roofline of the FMA kernel
(HIP events on the library's own stream, around the kernel alone), the
other binary256 operations' rates, and the reference's CPU path timed on this
host's cores on the same inputs, whose outputs are also compared with the
GPU's bit for bit.
"""

import argparse
import ctypes
import json
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
BYTES = {"add": 96, "mul": 96, "fma": 128, "div": 96, "sqrt": 64}


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
    return ap.parse_args()


def cpu_baseline(a, b, c, gpu_out, n, threads):
    """The reference's CPU path (its own C library, compiled from its sources
    into oracle/_ref by oracle/Makefile) on the same operands: one software
    device per thread over contiguous slices, every mode.  Falls back to the
    oracle restatement when the reference library is not present."""
    ref_path = os.path.join(HERE, "oracle", "_ref", "libcft_ref.so")
    port_path = os.path.join(HERE, "oracle", "_build", "libcftoracle.so")
    V, U32P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)
    if os.path.exists(ref_path):
        kind = "reference"
        lib = ctypes.CDLL(ref_path)
        lib.cft_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(V)]
        lib.cft_run.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                                ctypes.c_size_t, U32P, U32P]
        lib.cft_close.argtypes = [V]

        def run_slice(rnd, lo, hi, out):
            dev = V()
            assert lib.cft_open(None, 0, ctypes.byref(dev)) == 0
            f = ctypes.c_uint32()
            st = lib.cft_run(dev, 0, 3, rnd, a.ctypes.data + 32 * lo, b.ctypes.data + 32 * lo,
                             c.ctypes.data + 32 * lo, out.ctypes.data + 32 * lo, hi - lo,
                             ctypes.byref(f), None)
            lib.cft_close(dev)
            assert st == 0
    elif os.path.exists(port_path):
        kind = "port"
        lib = ctypes.CDLL(port_path)
        lib.oc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V, V, V,
                               ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, U32P]

        def run_slice(rnd, lo, hi, out):
            f = ctypes.c_uint32()
            lib.oc_run(3, 0, rnd, a.ctypes.data + 32 * lo, b.ctypes.data + 32 * lo,
                       c.ctypes.data + 32 * lo, out.ctypes.data + 32 * lo, hi - lo, 0, 1,
                       ctypes.byref(f))
    else:
        return None, None
    out = np.empty(n * 32, dtype=np.uint8)
    mismatches = 0
    t0 = time.perf_counter()
    for rnd in range(5):
        ths = []
        for t in range(threads):
            lo, hi = n * t // threads, n * (t + 1) // threads
            th = threading.Thread(target=run_slice, args=(rnd, lo, hi, out))
            th.start()
            ths.append(th)
        for th in ths:
            th.join()
        mismatches += int(np.count_nonzero(
            (out.reshape(-1, 32) != gpu_out[rnd][: n * 32].reshape(-1, 32)).any(axis=1)))
    secs = time.perf_counter() - t0
    return {"value": round(5 * n / secs / 1e9, 6), "unit": "Gelem/s", "cores": threads,
            "kind": kind,
            "sample": f"binary256 FMA, the first {n} elements of the benchmark's operands, "
                      f"all five rounding modes ({5 * n} elements), {threads} threads, one "
                      f"software device per thread, {secs:.1f} s wall"}, mismatches


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path shards by element index with no exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None

    from cftamd import Device, workload, _lib
    lib = _lib.lib()
    n = args.n
    log(f"bench: generating {n} binary256 operand triples (seed {args.seed})")
    t0 = time.perf_counter()
    dev = Device("hip:0")
    bufs = [dev.alloc(n * 32) for _ in range(4)]
    workload.elementwise(2, n, seed=args.seed, threads=16,
                         out=(bufs[0].host, bufs[1].host, bufs[2].host))
    for bf in bufs[:3]:
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
            lib.cftx_kernel_timer(dev.handle, 1)
            t1 = time.perf_counter()
            for rnd in range(5):
                if name == "add":
                    st = lib.cft_run(dev.handle, 1, 3, rnd, bufs[0].data_ptr, None,
                                     bufs[2].data_ptr, bufs[3].data_ptr, n, None, None)
                elif name == "mul":
                    st = lib.cft_run(dev.handle, 3, 3, rnd, bufs[0].data_ptr,
                                     bufs[1].data_ptr, None, bufs[3].data_ptr, n, None, None)
                elif name == "div":
                    st = lib.cft_div(dev.handle, 3, rnd, bufs[0].data_ptr, bufs[1].data_ptr,
                                     bufs[3].data_ptr, n, None, None)
                else:
                    st = lib.cft_sqrt(dev.handle, 3, rnd, bufs[0].data_ptr, bufs[3].data_ptr,
                                      n, None, None)
                assert st == 0, lib.cft_last_error()
            dt = time.perf_counter() - t1
            lib.cftx_kernel_timer_read(dev.handle, ctypes.byref(kms), ctypes.byref(klaunch))
            lib.cftx_kernel_timer(dev.handle, 0)
            ks = kms.value / 1e3 / max(1, klaunch.value)
            ops[name] = {"Gelem_s": round(5 * n / dt / 1e9, 4),
                         "kernel_ms": round(ks * 1e3, 4),
                         "kernel_GBs": round(BYTES[name] * n / ks / 1e9, 1)}
            log(f"bench: {name} {ops[name]}")

    cpu = None
    bit_exact = None
    if not args.skip_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        a = bufs[0].host.copy()
        b = bufs[1].host.copy()
        c = bufs[2].host.copy()
        log(f"bench: CPU baseline on {threads} threads")
        cpu, mism = cpu_baseline(a, b, c, gpu_out, n, threads)
        if cpu is not None:
            bit_exact = mism == 0
            log(f"bench: CPU {cpu['value']} Gelem/s ({cpu['kind']}), mismatches {mism}")

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
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel_ms": round(kavg_s * 1e3, 4),
                     "kernel_Gelem_s": round(n / kavg_s / 1e9, 4),
                     "bytes_per_elem": FMA_BYTES},
        "cpu_baseline": cpu,
        "bit_exact_vs_cpu": bit_exact,
        "ops": ops,
    }
    for bf in bufs:
        bf.free()
    dev.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
