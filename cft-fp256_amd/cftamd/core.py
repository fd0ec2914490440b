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
from ._lib import check, lib


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

    # ---- the 754 status word ------------------------------------------
    def save_all_flags(self):
        return lib().cft_save_all_flags(self.handle)

    def lower_flags(self, mask=0x1F):
        lib().cft_lower_flags(self.handle, mask)

    def alloc(self, nbytes):
        return Buffer(self, nbytes)


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
