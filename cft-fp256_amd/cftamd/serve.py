"""The serving loop from Python: a thin wrapper of cftx_server_*, and the
mixed request stream of the serving workload.

Workload (seeded): Poisson arrivals at `rate` requests/s for `duration`
seconds, each request one of
  * elementwise FMA / DIV / SQRT (60/20/20) in binary128 or binary256
    (50/50), round-to-nearest, n log-uniform in [1e3, 1e7] elements - a
    window of the operand pools (the elementwise workload's generators; SQRT
    reads |a|), the result into its own window;
  * Lorenz-63 RK4 binary256: 16 trajectories, steps log-uniform [1e3, 1e5],
    h = RN(1e-3), as the ensemble workload draws them;
  * Lorenz-96 RK4 binary256: one instance, ring K uniform in {8, ..., 256},
    steps uniform [1e3, 2e4];
  * Henon-Heiles leapfrog binary256: 16 orbits at E = 1/8, steps uniform
    [1e4, 1e5], h = RN(1e-2).
Operands, outputs and states live in resident buffers (device memory);
latency is submission (at the arrival instant) to completion as the server's
scheduler sees it.  Every job is checked afterwards against the same job run
alone (cft_run / cft_div / cft_sqrt / cftx_program_run_batch).
"""

import ctypes
import threading
import time

import numpy as np

from . import _lib, odes, workload
from . import programs as P
from ._lib import check, lib
from .core import DeviceArray, Program, run_multi

KIND_NAMES = {0: "fma", 1: "div", 2: "sqrt", 3: "lorenz63", 4: "lorenz96", 5: "henonheiles"}


class Server:
    """cftx_server on a Device's GPU."""

    def __init__(self, dev, workers=4, budget=0):
        self.dev = dev
        h = ctypes.c_void_p()
        check(lib().cftx_server_open(dev.handle, workers, budget, ctypes.byref(h)),
              "cftx_server_open")
        self.handle = h
        self._keep = {}

    def set_batching(self, max_jobs):
        check(lib().cftx_server_set_batching(self.handle, max_jobs), "cftx_server_set_batching")

    def set_slice(self, ms):
        check(lib().cftx_server_set_slice(self.handle, ms), "cftx_server_set_slice")

    def _submit(self, J, keep=None):
        J.struct_size = ctypes.sizeof(J)
        t = ctypes.c_uint64()
        check(lib().cftx_server_submit(self.handle, ctypes.byref(J), ctypes.byref(t)),
              "cftx_server_submit")
        return t.value

    def submit_elem(self, kind, op, fmt, rnd, a, b, c, d, n):
        J = _lib.CftxJob()
        J.kind, J.op, J.fmt, J.rnd = kind, op, fmt, rnd
        J.a, J.b, J.c, J.d, J.n = a, b, c, d, n
        return self._submit(J)

    def submit_program(self, prog, state, state_bytes, steps, bank=None):
        steps = np.ascontiguousarray(steps, dtype=np.uint32)
        J = _lib.CftxJob()
        J.kind = _lib.JOB_PROGRAM
        J.prog = prog.handle
        J.n = steps.size
        J.state, J.state_bytes = state, state_bytes
        J.steps = steps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        if bank is not None:
            J.bank, J.bank_bytes = bank.ctypes.data, bank.nbytes
        return self._submit(J)

    def wait(self, ticket):
        r = _lib.CftxJobResult()
        r.struct_size = ctypes.sizeof(r)
        check(lib().cftx_server_wait(self.handle, ticket, ctypes.byref(r)), "cftx_server_wait")
        return r

    def clock(self):
        return lib().cftx_server_clock(self.handle)

    def stats(self):
        s = _lib.CftxServerStats()
        s.struct_size = ctypes.sizeof(s)
        lib().cftx_server_stats_get(self.handle, ctypes.byref(s))
        return {f: getattr(s, f) for f, _t in s._fields_ if f != "struct_size"}

    def close(self):
        if self.handle:
            lib().cftx_server_close(self.handle)
            self.handle = None


def requests(seed, rate, duration, pool_n=10_000_000, mix=(0.8, 0.1, 0.05, 0.05)):
    """The seeded request list, sorted by arrival time (seconds)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        u = rng.random()
        if u < mix[0]:
            k = int(rng.choice([0, 0, 0, 1, 2]))
            fmt = int(rng.choice([2, 3]))
            n = int(np.clip(np.rint(np.exp(rng.uniform(np.log(1e3), np.log(1e7)))), 1000,
                            pool_n))
            off = int(rng.integers(0, pool_n - n + 1))
            out.append({"t": t, "kind": k, "fmt": fmt, "n": n, "off": off})
        elif u < mix[0] + mix[1]:
            out.append({"t": t, "kind": 3, "seed": int(rng.integers(1 << 31))})
        elif u < mix[0] + mix[1] + mix[2]:
            out.append({"t": t, "kind": 4, "seed": int(rng.integers(1 << 31)),
                        "k": int(rng.choice(np.arange(8, 257, 8)))})
        else:
            out.append({"t": t, "kind": 5, "seed": int(rng.integers(1 << 31))})
    return out


class Pools:
    """Device-resident operand pools, output windows and ODE programs."""

    def __init__(self, dev, reqs, seed=0, pool_n=10_000_000, verify_every=25,
                 device_out=False):
        self.dev, self.pool_n = dev, pool_n
        self.device_out = device_out
        self.ops = {}
        for fmt, cfg in ((2, 1), (3, 2)):
            esz = 16 if fmt == 2 else 32
            bufs = [dev.alloc(pool_n * esz) for _ in range(4)]
            workload.elementwise(cfg, pool_n, seed=seed, threads=16,
                                 out=(bufs[0].host, bufs[1].host, bufs[2].host))
            # sqrt reads |a|
            sa = bufs[3].host.reshape(pool_n, esz)
            sa[:] = bufs[0].host.reshape(pool_n, esz)
            sa[:, esz - 1] &= 0x7F
            for bf in bufs:
                bf.to_device()
            self.ops[fmt] = bufs
        # outputs: every `verify_every`-th elementwise job writes a window of
        # its own, kept for the bit check; the others share a sink.  With
        # device_out every job has a window of its own in raw device memory
        # (no pinned host mirror of the whole stream's output)
        if device_out:
            verify_every = 1
        need = {2: 0, 3: 0}
        for i, r in enumerate(reqs):
            if r["kind"] <= 2 and i % verify_every == 0:
                need[r["fmt"]] += r["n"]
        esz_of = {2: 16, 3: 32}
        if device_out:
            self.out = {f: DeviceArray(max(1, need[f]) * esz_of[f], fill=0xA5) for f in (2, 3)}
            self.sink = {}
        else:
            self.out = {f: dev.alloc(max(1, need[f]) * esz_of[f]) for f in (2, 3)}
            self.sink = {f: dev.alloc(pool_n * esz_of[f]) for f in (2, 3)}
            for f in (2, 3):
                self.out[f].to_device()
                self.sink[f].to_device()
        pos = {2: 0, 3: 0}
        for i, r in enumerate(reqs):
            if r["kind"] <= 2:
                r["keep"] = i % verify_every == 0
                if r["keep"]:
                    r["out"] = pos[r["fmt"]]
                    pos[r["fmt"]] += r["n"]
        # ODE programs, banks and states
        self.progs = {}
        self.banks = {}
        self.images = {"l63": P.with_trips(P.lorenz63(3), 100),
                       "hh": P.with_trips(P.henonheiles(3), 100)}
        self.progs["l63"] = Program(dev, self.images["l63"])
        self.banks["l63"] = np.frombuffer(P.bank_bytes("lorenz63-rk4", 3, "0.001"),
                                          dtype=np.uint8).copy()
        self.progs["hh"] = Program(dev, self.images["hh"])
        self.banks["hh"] = np.frombuffer(P.bank_bytes("henonheiles-lf", 3, "0.01"),
                                         dtype=np.uint8).copy()
        self.banks["l96"] = np.frombuffer(P.bank_bytes("lorenz96-rk4", 3, "0.01"),
                                          dtype=np.uint8).copy()
        states, total = [], 0
        for r in reqs:
            if r["kind"] == 3:
                _i, _b, st, steps = odes.lorenz63(16, seed=r["seed"], fmt=3)
            elif r["kind"] == 4:
                (_k, _i, _b, _ids, st, steps), = odes.lorenz96(1, seed=r["seed"], fmt=3,
                                                               ks=[r["k"]])
                if r["k"] not in self.progs:
                    self.images[r["k"]] = P.with_trips(P.lorenz96(3, k=r["k"]), 20)
                    self.progs[r["k"]] = Program(dev, self.images[r["k"]])
            elif r["kind"] == 5:
                _i, _b, st, steps = odes.henonheiles(16, seed=r["seed"], fmt=3)
            else:
                continue
            r["state_off"], r["state_bytes"], r["steps"] = total, st.nbytes, steps
            states.append(st)
            total += st.nbytes
        for p in self.progs.values():          # schedules built ahead of serving
            p.prepare(0x7F)
        self.state0 = np.concatenate(states) if states else np.zeros(32, np.uint8)
        self.states = dev.alloc(max(32, total))
        self.states.host[: self.state0.nbytes] = self.state0
        self.states.to_device()
        self.verify_every = verify_every

    def prog_of(self, r):
        return {3: self.progs["l63"], 5: self.progs["hh"]}.get(r["kind"]) or self.progs[r["k"]]

    def image_of(self, r):
        return {3: self.images["l63"], 5: self.images["hh"]}.get(r["kind"]) or self.images[r["k"]]

    def bank_of(self, r):
        return {3: self.banks["l63"], 4: self.banks["l96"], 5: self.banks["hh"]}[r["kind"]]

    def out_ptr(self, fmt):
        o = self.out[fmt]
        return o.ptr if self.device_out else o.data_ptr

    def served(self, r):
        """A kept elementwise job's output bytes (after the stream ran; for
        resident outputs after their from_device)."""
        esz = 16 if r["fmt"] == 2 else 32
        lo, nb = r["out"] * esz, r["n"] * esz
        if self.device_out:
            return self.out[r["fmt"]].download(lo, nb)
        return self.out[r["fmt"]].host[lo:lo + nb]

    def free(self):
        for bufs in self.ops.values():
            for b in bufs:
                b.free()
        for b in list(self.out.values()) + list(self.sink.values()) + [self.states]:
            b.free()
        for p in self.progs.values():
            p.free()


def submit(srv, pools, r):
    if r["kind"] <= 2:
        fmt, esz = r["fmt"], (16 if r["fmt"] == 2 else 32)
        a, b, c, sa = (x.data_ptr + r["off"] * esz for x in pools.ops[fmt])
        d = (pools.out_ptr(fmt) + r["out"] * esz if r["keep"] else
             pools.sink[fmt].data_ptr)
        if r["kind"] == 0:
            return srv.submit_elem(_lib.JOB_ELEM, 0, fmt, 0, a, b, c, d, r["n"])
        if r["kind"] == 1:
            return srv.submit_elem(_lib.JOB_DIV, 0, fmt, 0, a, b, None, d, r["n"])
        return srv.submit_elem(_lib.JOB_SQRT, 0, fmt, 0, sa, None, None, d, r["n"])
    return srv.submit_program(pools.prog_of(r), pools.states.data_ptr + r["state_off"],
                              r["state_bytes"], r["steps"], bank=pools.bank_of(r))


def serve(srv, pools, reqs):
    """Submit every request at its arrival time (real time) and collect the
    results; returns per-request results and the wall time."""
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()
    done = threading.Event()

    def feeder():
        for i, r in enumerate(reqs):
            delay = r["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = submit(srv, pools, r)
        done.set()

    th = threading.Thread(target=feeder)
    th.start()
    results = []
    for i in range(len(reqs)):
        while tickets[i] is None:
            time.sleep(0.0005)
        results.append(srv.wait(tickets[i]))
    th.join()
    return results, time.perf_counter() - t0


def summarize(reqs, results, wall):
    lat = np.array([r.done_s - r.submit_s for r in results])
    out = {"requests": len(reqs), "wall_s": round(wall, 3),
           "throughput_req_s": round(len(reqs) / wall, 1),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "by_kind": {}}
    for k, name in KIND_NAMES.items():
        sel = np.array([r["kind"] == k for r in reqs])
        if sel.any():
            out["by_kind"][name] = {"n": int(sel.sum()),
                                    "p50_ms": round(float(np.percentile(lat[sel], 50)) * 1e3, 3),
                                    "p99_ms": round(float(np.percentile(lat[sel], 99)) * 1e3, 3)}
    elem = sum(r["n"] for r in reqs if r["kind"] <= 2)
    out["elements"] = int(elem)
    out["ode_steps"] = int(sum(int(np.asarray(r["steps"]).sum()) for r in reqs if r["kind"] >= 3))
    out["statuses_ok"] = all(r.status == 0 for r in results)
    return out


def verify(dev, pools, reqs, results, alone_every=1, hook=None):
    """Every job against the same job run on its own: flags for every
    elementwise job (cft_run / cft_div / cft_sqrt), bits for every kept one
    (every `verify_every`-th, or all of them with device_out); final states,
    flags and STATUS for every program job.  Every `alone_every`-th program
    job is rerun by itself (cftx_program_run_batch); the others in one
    cftx_program_run_multi launch, a part each (a lane's results depend on
    its own state and trips only).  hook(r, result, served_bytes) sees every
    kept elementwise job's output."""
    L = lib()
    if not pools.device_out:
        for b in pools.out.values():
            b.from_device()
    pools.states.from_device()
    served_states = pools.states.host.copy()
    scratch = {f: dev.alloc(pools.pool_n * (16 if f == 2 else 32)) for f in (2, 3)}
    bad = checked = bits = alone = 0
    fl = ctypes.c_uint32()
    multi = []
    for i, (r, res) in enumerate(zip(reqs, results)):
        checked += 1
        if r["kind"] <= 2:
            fmt, esz = r["fmt"], (16 if r["fmt"] == 2 else 32)
            a, b, c, sa = (x.data_ptr + r["off"] * esz for x in pools.ops[fmt])
            # a kept job's result comes back to the host, the others stay on the device
            host = np.empty(r["n"] * esz, np.uint8) if r["keep"] else None
            d = host.ctypes.data if r["keep"] else scratch[fmt].data_ptr
            if r["kind"] == 0:
                st = L.cft_run(dev.handle, 0, fmt, 0, a, b, c, d, r["n"], ctypes.byref(fl), None)
            elif r["kind"] == 1:
                st = L.cft_div(dev.handle, fmt, 0, a, b, d, r["n"], ctypes.byref(fl), None)
            else:
                st = L.cft_sqrt(dev.handle, fmt, 0, sa, d, r["n"], ctypes.byref(fl), None)
            ok = st == 0 and fl.value == res.flags
            if r["keep"]:
                served = pools.served(r)
                if hook is not None:
                    hook(r, res, served)
                ok = ok and np.array_equal(host, served)
                bits += 1
            bad += 0 if ok else 1
            continue
        lo, hi = r["state_off"], r["state_off"] + r["state_bytes"]
        st = pools.state0[lo:hi].copy()
        if alone_every and i % alone_every == 0:
            f, s = pools.prog_of(r).run_batch(st, r["steps"], bank=pools.bank_of(r))
            alone += 1
            ok = np.array_equal(st, served_states[lo:hi]) and (f, s) == (res.flags, res.bus)
            bad += 0 if ok else 1
        else:
            lf = np.zeros(np.asarray(r["steps"]).size, np.uint8)
            multi.append((r, res, st, lf))
    if multi:
        parts = [dict(prog=pools.prog_of(r), bank=pools.bank_of(r), state=st, steps=r["steps"],
                      lane_flags=lf) for r, _res, st, lf in multi]
        run_multi(dev, parts)
        for r, res, st, lf in multi:
            lo, hi = r["state_off"], r["state_off"] + r["state_bytes"]
            f = int(np.bitwise_or.reduce(lf & 0x1F)) if lf.size else 0
            ok = np.array_equal(st, served_states[lo:hi]) and f == res.flags and res.bus == 0
            bad += 0 if ok else 1
    for b in scratch.values():
        b.free()
    return {"checked": checked, "bits_checked_elementwise": bits, "programs_alone": alone,
            "programs_in_one_launch": len(multi), "mismatches": bad}
