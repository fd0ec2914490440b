"""Devices, elementwise runs and resident buffers over the C ABI.

Operands are numpy uint8 arrays of dense little-endian encodings (n * esz
bytes), or resident :class:`Buffer` objects, exactly what the C calls take.
The call shapes follow the reference binding's batch API (bindings/python/
cftmpfr/batch.py:145-196: one C call per op) and the argument rules of
cft_run (host/include/cft.h:1166-1209).
"""

import ctypes
from ctypes import byref, c_uint32, c_void_p

import numpy as np

from . import _lib
from ._lib import CftError, check, lib


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, Buffer):
        return x.data_ptr
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("operands must be C-contiguous")
        return x.ctypes.data
    raise TypeError(f"operand of type {type(x).__name__}")


class Device:
    """A cft_device on MI355X: ``Device("hip:0")``."""

    def __init__(self, artifact="hip:0", index=0, scratch_depth=None):
        h = c_void_p()
        if scratch_depth is None:
            check(lib().cft_open(artifact.encode() if artifact else None, index,
                                 byref(h)), "cft_open")
        else:
            args = _lib.CftOpenArgs(ctypes.sizeof(_lib.CftOpenArgs),
                                    artifact.encode() if artifact else None,
                                    index, scratch_depth)
            check(lib().cft_open_ex(byref(args), byref(h)), "cft_open_ex")
        self.handle = h

    def close(self):
        if self.handle:
            lib().cft_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def caps(self):
        c = _lib.CftCaps()
        c.struct_size = ctypes.sizeof(c)
        check(lib().cft_get_caps(self.handle, byref(c)), "cft_get_caps")
        return c

    # ---- elementwise --------------------------------------------------
    def run(self, op, fmt, rnd, a, b=None, c=None, out=None, n=None):
        """d = op(a, b, c) elementwise; returns (d, flags)."""
        esz = _lib.FORMAT_BYTES[fmt]
        if n is None:
            src = a if a is not None else (b if b is not None else c)
            n = (src.nbytes if isinstance(src, np.ndarray) else src.bytes) // esz
        if out is None:
            out = np.empty(n * esz, dtype=np.uint8)
        fl = c_uint32()
        bus = c_uint32()
        check(lib().cft_run(self.handle, op, fmt, rnd, _ptr(a), _ptr(b), _ptr(c),
                            _ptr(out), n, byref(fl), byref(bus)), "cft_run")
        return out, fl.value

    def div(self, fmt, rnd, a, b, out=None, n=None):
        esz = _lib.FORMAT_BYTES[fmt]
        n = n if n is not None else (a.nbytes if isinstance(a, np.ndarray) else a.bytes) // esz
        if out is None:
            out = np.empty(n * esz, dtype=np.uint8)
        fl = c_uint32()
        check(lib().cft_div(self.handle, fmt, rnd, _ptr(a), _ptr(b), _ptr(out), n,
                            byref(fl), None), "cft_div")
        return out, fl.value

    def sqrt(self, fmt, rnd, a, out=None, n=None):
        esz = _lib.FORMAT_BYTES[fmt]
        n = n if n is not None else (a.nbytes if isinstance(a, np.ndarray) else a.bytes) // esz
        if out is None:
            out = np.empty(n * esz, dtype=np.uint8)
        fl = c_uint32()
        check(lib().cft_sqrt(self.handle, fmt, rnd, _ptr(a), _ptr(out), n, byref(fl),
                             None), "cft_sqrt")
        return out, fl.value

    # ---- reductions (host/include/cft.h:1210-1340) ---------------------
    def reduce(self, op, fmt, rnd, a, b=None, out=None, n=None, seg=None):
        """The reduction tree over a (and b for CFT_DOT): one result, or one
        per segment of `seg` elements; returns (d, flags)."""
        esz = _lib.FORMAT_BYTES.get(fmt, 1)          # the library refuses a bad format
        if n is None:
            n = (a.nbytes if isinstance(a, np.ndarray) else a.bytes) // esz
        nres = 1 if seg is None else (n // seg if seg else 0)
        if out is None:
            out = np.zeros(max(nres, 1) * esz, dtype=np.uint8)
        fl = c_uint32()
        if seg is None:
            check(lib().cft_reduce(self.handle, op, fmt, rnd, _ptr(a), _ptr(b), _ptr(out), n,
                                   byref(fl), None), "cft_reduce")
        else:
            check(lib().cft_reduce_seg(self.handle, op, fmt, rnd, _ptr(a), _ptr(b), _ptr(out),
                                       n, seg, byref(fl), None), "cft_reduce_seg")
        return out, fl.value

    # ---- the 754 status word ------------------------------------------
    def save_all_flags(self):
        return lib().cft_save_all_flags(self.handle)

    def lower_flags(self, mask=0x1F):
        lib().cft_lower_flags(self.handle, mask)

    def alloc(self, nbytes):
        return Buffer(self, nbytes)

    def load(self, image):
        """A program from its image bytes (cft_program_load)."""
        return Program(self, image)


def _u32p(x):
    if x is None:
        return None
    p = _ptr(x)
    return ctypes.cast(p, ctypes.POINTER(c_uint32))


class Program:
    """A loaded sequencer program (cft_program_load) and its runs
    (cft_program_run_ex), with the run's arrays as numpy uint8 / uint32
    arrays or resident buffers, laid out as cft_run_args has them."""

    def __init__(self, dev, image):
        self.dev = dev
        self._image = np.frombuffer(bytes(image), dtype=np.uint8).copy()
        h = c_void_p()
        check(lib().cft_program_load(dev.handle, self._image.ctypes.data, self._image.nbytes,
                                     byref(h)), "cft_program_load")
        self.handle = h

    def info(self):
        i = _lib.CftProgramInfo()
        i.struct_size = ctypes.sizeof(i)
        check(lib().cft_program_get_info(self.handle, byref(i)), "cft_program_get_info")
        return i

    def digest(self, bank=None):
        out = (ctypes.c_uint8 * 32)()
        nb = 0 if bank is None else (bank.nbytes if isinstance(bank, np.ndarray) else bank.bytes)
        check(lib().cft_program_digest(self.handle, _ptr(bank), nb, out), "cft_program_digest")
        return bytes(out)

    def run(self, n, a, b=None, c=None, bank=None, scratch_in=None, scratch_out=None,
            deposits=None, counts=None, lane_mask=None, lane_flags=None, idx_a=None,
            idx_b=None, idx_c=None, idx_a_src=0, idx_b_src=0, idx_c_src=0,
            idx_scratch_in=None, idx_scratch_src=0):
        """One cft_program_run_ex over n lanes; returns (flags, status)."""
        def nbytes(x):
            if x is None:
                return 0
            return x.nbytes if isinstance(x, np.ndarray) else x.bytes
        A = _lib.CftRunArgs()
        A.struct_size = ctypes.sizeof(A)
        A.a, A.b, A.c, A.n = _ptr(a), _ptr(b), _ptr(c), n
        A.bank, A.bank_bytes = _ptr(bank), nbytes(bank)
        A.scratch_in, A.scratch_in_bytes = _ptr(scratch_in), nbytes(scratch_in)
        A.scratch_out, A.scratch_out_bytes = _ptr(scratch_out), nbytes(scratch_out)
        A.deposits = _ptr(deposits)
        A.counts = _u32p(counts)
        fl, bus = c_uint32(), c_uint32()
        A.flags_out = ctypes.pointer(fl)
        A.bus_out = ctypes.pointer(bus)
        A.idx_a, A.idx_b, A.idx_c = _u32p(idx_a), _u32p(idx_b), _u32p(idx_c)
        A.idx_a_src, A.idx_b_src, A.idx_c_src = idx_a_src, idx_b_src, idx_c_src
        A.idx_scratch_in, A.idx_scratch_src = _u32p(idx_scratch_in), idx_scratch_src
        A.lane_mask, A.lane_mask_bytes = _ptr(lane_mask), nbytes(lane_mask)
        A.lane_flags, A.lane_flags_bytes = _ptr(lane_flags), nbytes(lane_flags)
        check(lib().cft_program_run_ex(self.handle, byref(A)), "cft_program_run_ex")
        return fl.value, bus.value

    def run_batch(self, state, steps, bank=None, a=None, b=None, c=None, lane_flags=None,
                  mode=_lib.BATCH_PERSISTENT, lanes_per_wave=0, waves_per_group=0,
                  graph=False, lanes_per_instance=0):
        """cftx_program_run_batch: lane i runs steps[i] trips of the segment
        loop, its state (n x slots elements) updated in place; returns
        (flags, status)."""
        steps = np.ascontiguousarray(steps, dtype=np.uint32)
        n = steps.size
        nb = lambda x: 0 if x is None else (x.nbytes if isinstance(x, np.ndarray) else x.bytes)
        A = _lib.CftxBatchArgs()
        A.struct_size = ctypes.sizeof(A)
        A.bank, A.bank_bytes = _ptr(bank), nb(bank)
        A.state, A.state_bytes = _ptr(state), nb(state)
        A.steps = steps.ctypes.data_as(ctypes.POINTER(c_uint32))
        A.n = n
        A.a, A.b, A.c = _ptr(a), _ptr(b), _ptr(c)
        fl, bus = c_uint32(), c_uint32()
        A.flags_out = ctypes.pointer(fl)
        A.bus_out = ctypes.pointer(bus)
        A.lane_flags, A.lane_flags_bytes = _ptr(lane_flags), nb(lane_flags)
        A.mode = mode
        A.lanes_per_wave = lanes_per_wave
        A.waves_per_group = waves_per_group
        A.graph = 1 if graph else 0
        A.lanes_per_instance = lanes_per_instance
        check(lib().cftx_program_run_batch(self.handle, byref(A)), "cftx_program_run_batch")
        return fl.value, bus.value

    def prepare(self, widths=0x7F):
        """cftx_program_prepare: build the wide schedules named by the bit
        mask (bit k: 2^k lanes an instance); returns the mask that built."""
        ready = c_uint32()
        check(lib().cftx_program_prepare(self.handle, widths, byref(ready)),
              "cftx_program_prepare")
        return ready.value

    def free(self):
        if self.handle:
            lib().cft_program_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def run_multi(dev, parts, waves_per_group=0, waves_per_cu=0, alone=0.0):
    """cftx_program_run_multi over parts, each a dict with prog (a Program),
    state (updated in place), steps, and optionally bank, lane_flags and
    lanes_per_instance; the planner's knobs after.  Returns (flags, status,
    widths used)."""
    arr = (_lib.CftxBatchPart * len(parts))()
    keep = []
    nb = lambda x: 0 if x is None else (x.nbytes if isinstance(x, np.ndarray) else x.bytes)
    for k, p in enumerate(parts):
        steps = np.ascontiguousarray(p["steps"], dtype=np.uint32)
        keep.append(steps)
        P = arr[k]
        P.struct_size = ctypes.sizeof(P)
        P.prog = p["prog"].handle
        P.bank, P.bank_bytes = _ptr(p.get("bank")), nb(p.get("bank"))
        P.state, P.state_bytes = _ptr(p["state"]), nb(p["state"])
        P.steps = steps.ctypes.data_as(ctypes.POINTER(c_uint32))
        P.n = steps.size
        P.lane_flags = _ptr(p.get("lane_flags"))
        P.lanes_per_instance = p.get("lanes_per_instance", 0)
    fl, bus = c_uint32(), c_uint32()
    opts = _lib.CftxMultiOpts(ctypes.sizeof(_lib.CftxMultiOpts), waves_per_group, waves_per_cu,
                              alone)
    check(lib().cftx_program_run_multi(dev.handle, arr, len(parts), byref(opts),
                                       byref(fl), byref(bus)), "cftx_program_run_multi")
    return fl.value, bus.value, [arr[k].width_used for k in range(len(parts))]


def last_plan(dev):
    """cftx_multi_last_plan: the handle's last multi-program launch plan."""
    out = (ctypes.c_uint64 * 6)()
    check(lib().cftx_multi_last_plan(dev.handle, out), "cftx_multi_last_plan")
    keys = ("waves_per_group", "groups", "lds_rows_per_wave", "tiles", "groups_per_cu",
            "planned_bundles")
    return dict(zip(keys, [int(v) for v in out]))


def tune_multi(dev, parts, trials=None):
    """The autotuner for cftx_program_run_multi: each setting in `trials`
    runs on copies of the parts' states and the fastest is returned with
    every trial's seconds.  A setting is a dict of run_multi's knobs
    (alone, waves_per_cu, waves_per_group) and optionally "width": lanes an
    instance forced on every part (instance packing: 64 / width instances a
    wave).  By default: the planner's lone-wave fraction 0.3..1.0; 4 and 8
    waves a CU; workgroups of 1, 2 and 4 waves; and every part at 1, 2, 4,
    8 and 16 lanes an instance.  Run it on a sample of the workload (steps
    scaled down): the choices depend on the ratios of the steps, not their
    size."""
    import time
    if trials is None:
        trials = [dict(alone=a) for a in (0.3, 0.5, 0.7, 0.9, 1.0)]
        trials += [dict(alone=a, waves_per_cu=k) for a in (0.5, 0.9) for k in (4, 8)]
        trials += [dict(waves_per_group=g) for g in (1, 2, 4)]
        trials += [dict(width=w) for w in (1, 2, 4, 8, 16)]
    results = []
    for t in trials:
        knobs = {k: v for k, v in t.items() if k != "width"}
        trial = [dict(p, state=np.array(p["state"], copy=True)) for p in parts]
        if "width" in t:
            for q in trial:
                q["lanes_per_instance"] = t["width"]
        try:
            t0 = time.perf_counter()
            run_multi(dev, trial, **knobs)
            results.append((time.perf_counter() - t0, t))
        except CftError:
            continue
    results.sort(key=lambda r: r[0])
    return results[0][1], [(round(s, 5), t) for s, t in results]


def apply_knobs(parts, knobs):
    """run_multi's keyword arguments for a tune_multi setting, and the parts
    with its forced width (if any) applied."""
    kw = {k: v for k, v in knobs.items() if k != "width"}
    if "width" in knobs:
        parts = [dict(p, lanes_per_instance=knobs["width"]) for p in parts]
    return parts, kw


def tune_batch(prog, state, steps, bank=None, trials=None, repeats=1):
    """The autotuner for one program's batched run (cftx_program_run_batch,
    persistent): each setting - lanes an instance (1 interprets an instance
    on one lane; 2..64 spread it, pure programs only; 0 the library's model)
    and workgroup size in waves - runs on a copy of the state `repeats`
    times, and the setting with the fastest run is returned with every
    trial's best seconds.  Run it on a sample (steps scaled down)."""
    import time
    if trials is None:
        trials = [dict(lanes_per_instance=w, waves_per_group=g)
                  for w in (0, 1, 2, 4, 8, 16) for g in (0, 1, 2, 4)]
    results = []
    for t in trials:
        best = None
        for _ in range(max(1, repeats)):
            st = np.array(state, copy=True)
            try:
                t0 = time.perf_counter()
                prog.run_batch(st, steps, bank=bank, **t)
                dt = time.perf_counter() - t0
            except CftError:
                break
            best = dt if best is None else min(best, dt)
        if best is not None:
            results.append((best, t))
    results.sort(key=lambda r: r[0])
    return results[0][1], [(round(s, 5), t) for s, t in results]


def check_image(image, scratch_depth=0):
    """The loader's verdict on an image without a device (cftx_program_check):
    (status, cft_program_info or None)."""
    buf = np.frombuffer(bytes(image), dtype=np.uint8).copy()
    i = _lib.CftProgramInfo()
    i.struct_size = ctypes.sizeof(i)
    st = lib().cftx_program_check(buf.ctypes.data if buf.nbytes else None, buf.nbytes,
                                  scratch_depth, byref(i))
    return st, (i if st == _lib.OK else None)


class Buffer:
    """A resident buffer: a pinned host mirror (``.host``, a numpy view) and a
    device copy that runs read and write in place."""

    def __init__(self, dev, nbytes):
        h = c_void_p()
        check(lib().cft_alloc(dev.handle, nbytes, byref(h)), "cft_alloc")
        self.dev = dev
        self.handle = h
        self.bytes = nbytes
        self.data_ptr = lib().cft_buffer_data(h)
        self.host = np.ctypeslib.as_array(
            ctypes.cast(self.data_ptr, ctypes.POINTER(ctypes.c_uint8)), shape=(nbytes,))

    def to_device(self):
        check(lib().cft_buffer_to_device(self.handle), "cft_buffer_to_device")

    def from_device(self):
        check(lib().cft_buffer_from_device(self.handle), "cft_buffer_from_device")

    def info(self):
        i = _lib.CftBufferInfo()
        i.struct_size = ctypes.sizeof(i)
        check(lib().cft_buffer_get_info(self.handle, byref(i)), "cft_buffer_get_info")
        return i

    def free(self):
        if self.handle:
            lib().cft_buffer_free(self.handle)
            self.handle = None
            self.host = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """Raw device memory of the handle's GPU (hipMalloc of the HIP runtime
    libcft itself is linked against), for job windows that need no host
    mirror: the serving loop takes such windows as well as resident ones.
    ``ptr`` is the device address; ``download`` copies bytes back."""

    _rt = None

    @classmethod
    def runtime(cls):
        if cls._rt is None:
            lib()                                 # libcft loads the runtime first
            rt = ctypes.CDLL("libamdhip64.so.7")
            rt.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), ctypes.c_size_t]
            rt.hipFree.argtypes = [c_void_p]
            rt.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, ctypes.c_int]
            rt.hipMemset.argtypes = [c_void_p, ctypes.c_int, ctypes.c_size_t]
            cls._rt = rt
        return cls._rt

    def __init__(self, nbytes, fill=None):
        rt = self.runtime()
        p = c_void_p()
        st = rt.hipMalloc(byref(p), max(1, nbytes))
        if st != 0:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed ({st})")
        self.ptr, self.nbytes = p.value, nbytes
        if fill is not None and rt.hipMemset(self.ptr, fill, nbytes) != 0:
            raise RuntimeError("hipMemset failed")

    def download(self, offset=0, nbytes=None, out=None):
        nbytes = self.nbytes - offset if nbytes is None else nbytes
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("window outside the allocation")
        if out is None:
            out = np.empty(nbytes, np.uint8)
        if self.runtime().hipMemcpy(out.ctypes.data, self.ptr + offset, nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy device to host failed")
        return out

    def free(self):
        if self.ptr:
            self.runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
