"""Loading and binding libcft.so - the MI355X engine's C ABI - through ctypes.

Mirrors the conventions of the reference binding (bindings/python/cftmpfr/
_lib.py:147-410): ``CFT_LIB`` names the library explicitly and wins outright;
otherwise the in-tree build at ``cft-fp256_amd/lib/libcft.so`` is used.  After
binding, ``_audit`` checks that the library's ABI major is 0 and that every
opcode and format number below means what its name says.  A failed call
raises :class:`CftError` carrying ``cft_last_error()``.

There is deliberately no fallback: if the HIP library is not there, loading
fails loudly.
"""

import ctypes
import os
from ctypes import POINTER, c_char_p, c_int, c_size_t, c_uint32, c_void_p

ABI_MAJOR = 0

# cft_status (host/include/cft.h:244-267)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_UNSUPPORTED = 2
ERR_NO_DEVICE = 3
ERR_ARTIFACT = 4
ERR_BUS_FAULT = 5
ERR_OUT_OF_MEMORY = 6
ERR_TIMEOUT = 7
ERR_INTERNAL = 8

# cft_format
FP32, FP64, FP128, FP256 = 0, 1, 2, 3
FORMAT_BYTES = {FP32: 4, FP64: 8, FP128: 16, FP256: 32}

# cft_op (host/include/cft.h:325-472)
OP_FMA, OP_ADD, OP_SUB, OP_MUL = 0, 1, 2, 3
OP_ABS, OP_NEG, OP_COPYSIGN = 4, 5, 6
OP_MIN, OP_MAX, OP_MINNUM, OP_MAXNUM = 7, 8, 9, 10
OP_SELECT, OP_CMPLT, OP_CMPLE, OP_CMPEQ = 11, 12, 13, 14
OP_IAND, OP_IOR, OP_IXOR, OP_IADD = 16, 17, 18, 19
OP_ISUB, OP_ISHL, OP_ISHR, OP_ICMPLT = 20, 21, 22, 23
OP_RECIP_SEED, OP_RSQRT_SEED = 26, 27
OP_IMUL = 30

# cft_round
RNE, RTZ, RDN, RUP, RMM = 0, 1, 2, 3, 4

# cft_exception
FLAG_INVALID, FLAG_DIVBYZERO, FLAG_OVERFLOW = 1, 2, 4
FLAG_UNDERFLOW, FLAG_INEXACT = 8, 16


class CftError(RuntimeError):
    """A libcft call returned a status other than CFT_OK."""

    def __init__(self, status, where, detail):
        self.status = status
        super().__init__(f"{where}: status {status}: {detail}")


class CftCaps(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("format_mask", c_uint32),
                ("tiles", c_uint32), ("abi_version", c_uint32),
                ("device_version", c_uint32), ("flags_readable", c_int),
                ("backend", ctypes.c_char * 32), ("max_deposits", c_uint32),
                ("max_insns", c_uint32), ("max_consts", c_uint32),
                ("seq_features", c_uint32), ("max_scratch", c_uint32),
                ("buffers_resident", c_int)]


class CftOpenArgs(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("artifact", c_char_p),
                ("index", c_int), ("scratch_depth", c_uint32)]


class CftBufferInfo(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("bytes", c_size_t),
                ("resident", c_int), ("device_authority", c_int),
                ("resident_binds", ctypes.c_uint64),
                ("staged_binds", ctypes.c_uint64),
                ("staged_why", ctypes.c_char * 112)]


class CftProgramInfo(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("format", c_int),
                ("max_deposits", c_uint32), ("n_insns", c_uint32),
                ("n_consts", c_uint32), ("flags", c_uint32),
                ("n_scratch_in", c_uint32), ("n_scratch_out", c_uint32),
                ("scratch_used", c_uint32)]


class CftRunArgs(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("a", c_void_p), ("b", c_void_p),
                ("c", c_void_p), ("n", c_size_t),
                ("bank", c_void_p), ("bank_bytes", c_size_t),
                ("scratch_in", c_void_p), ("scratch_in_bytes", c_size_t),
                ("scratch_out", c_void_p), ("scratch_out_bytes", c_size_t),
                ("deposits", c_void_p), ("counts", POINTER(c_uint32)),
                ("flags_out", POINTER(c_uint32)), ("bus_out", POINTER(c_uint32)),
                ("idx_a", POINTER(c_uint32)), ("idx_b", POINTER(c_uint32)),
                ("idx_c", POINTER(c_uint32)),
                ("idx_a_src", c_size_t), ("idx_b_src", c_size_t),
                ("idx_c_src", c_size_t),
                ("idx_scratch_in", POINTER(c_uint32)),
                ("idx_scratch_src", c_size_t),
                ("lane_mask", c_void_p), ("lane_mask_bytes", c_size_t),
                ("lane_flags", c_void_p), ("lane_flags_bytes", c_size_t)]


class CftElemArgs(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("a", c_void_p), ("b", c_void_p),
                ("c", c_void_p), ("d", c_void_p), ("n", c_size_t),
                ("scalar_mask", c_uint32), ("flags_out", POINTER(c_uint32)),
                ("bus_out", POINTER(c_uint32)),
                ("idx_a", POINTER(c_uint32)), ("idx_b", POINTER(c_uint32)),
                ("idx_c", POINTER(c_uint32)),
                ("idx_a_src", c_size_t), ("idx_b_src", c_size_t),
                ("idx_c_src", c_size_t)]


class CftxBatchArgs(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("bank", c_void_p), ("bank_bytes", c_size_t),
                ("state", c_void_p), ("state_bytes", c_size_t),
                ("steps", POINTER(c_uint32)), ("n", c_size_t),
                ("a", c_void_p), ("b", c_void_p), ("c", c_void_p),
                ("flags_out", POINTER(c_uint32)), ("bus_out", POINTER(c_uint32)),
                ("lane_flags", c_void_p), ("lane_flags_bytes", c_size_t),
                ("mode", c_int), ("lanes_per_wave", c_uint32),
                ("waves_per_group", c_uint32), ("graph", c_int),
                ("lanes_per_instance", c_uint32)]


BATCH_PERSISTENT, BATCH_PER_OP = 0, 1


class CftxMultiOpts(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("waves_per_group", c_uint32),
                ("waves_per_cu", c_uint32), ("alone", ctypes.c_double)]


class CftxBatchPart(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("prog", c_void_p),
                ("bank", c_void_p), ("bank_bytes", c_size_t),
                ("state", c_void_p), ("state_bytes", c_size_t),
                ("steps", POINTER(c_uint32)), ("n", c_size_t),
                ("lane_flags", c_void_p), ("lanes_per_instance", c_uint32),
                ("width_used", c_uint32)]

JOB_ELEM, JOB_DIV, JOB_SQRT, JOB_PROGRAM = 0, 1, 2, 3


class CftxJob(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("kind", c_int), ("op", c_int), ("fmt", c_int),
                ("rnd", c_int), ("a", c_void_p), ("b", c_void_p), ("c", c_void_p),
                ("d", c_void_p), ("n", c_size_t), ("prog", c_void_p),
                ("bank", c_void_p), ("bank_bytes", c_size_t),
                ("state", c_void_p), ("state_bytes", c_size_t),
                ("steps", POINTER(c_uint32))]


class CftxJobResult(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("status", c_int), ("flags", c_uint32),
                ("bus", c_uint32), ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("batch_jobs", c_uint32)]


class CftxServerStats(ctypes.Structure):
    _fields_ = [("struct_size", c_size_t), ("jobs", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("elem_batches", ctypes.c_uint64),
                ("program_batches", ctypes.c_uint64), ("max_batch_jobs", ctypes.c_uint64),
                ("peak_bytes", ctypes.c_uint64), ("slices", ctypes.c_uint64),
                ("max_slice_lanes", ctypes.c_uint64), ("slice_host_s", ctypes.c_double)]

HERE = os.path.dirname(os.path.abspath(__file__))
IN_TREE = os.path.join(os.path.dirname(HERE), "lib", "libcft.so")


def find_library():
    """CFT_LIB, else the in-tree build.  Raises if neither exists."""
    override = os.environ.get("CFT_LIB")
    if override:
        return override
    if os.path.exists(IN_TREE):
        return IN_TREE
    raise OSError(
        f"libcft.so is not built: expected {IN_TREE}. Run "
        f"`make -C cft-fp256_amd` (or __graft_entry__.build()), or point "
        f"CFT_LIB at the library.")


def _bind(lib):
    P = POINTER(c_uint32)
    V = c_void_p
    sig = {
        "cft_abi_version": ([], c_uint32),
        "cft_strerror": ([c_int], c_char_p),
        "cft_last_error": ([], c_char_p),
        "cft_format_size": ([c_int], c_size_t),
        "cft_format_name": ([c_int], c_char_p),
        "cft_op_name": ([c_int], c_char_p),
        "cft_open": ([c_char_p, c_int, POINTER(c_void_p)], c_int),
        "cft_open_ex": ([POINTER(CftOpenArgs), POINTER(c_void_p)], c_int),
        "cft_close": ([V], None),
        "cft_get_caps": ([V, POINTER(CftCaps)], c_int),
        "cft_supports": ([V, c_int, c_int], c_int),
        "cft_run": ([V, c_int, c_int, c_int, V, V, V, V, c_size_t, P, P], c_int),
        "cft_run_ex": ([V, c_int, c_int, c_int, POINTER(CftElemArgs)], c_int),
        "cft_div": ([V, c_int, c_int, V, V, V, c_size_t, P, P], c_int),
        "cft_reduce": ([V, c_int, c_int, c_int, V, V, V, c_size_t, P, P], c_int),
        "cft_reduce_seg": ([V, c_int, c_int, c_int, V, V, V, c_size_t, c_size_t, P, P],
                           c_int),
        "cft_sqrt": ([V, c_int, c_int, V, V, c_size_t, P, P], c_int),
        "cft_alloc": ([V, c_size_t, POINTER(c_void_p)], c_int),
        "cft_buffer_data": ([V], c_void_p),
        "cft_buffer_to_device": ([V], c_int),
        "cft_buffer_from_device": ([V], c_int),
        "cft_buffer_free": ([V], None),
        "cft_buffer_get_info": ([V, POINTER(CftBufferInfo)], c_int),
        "cft_program_load": ([V, V, c_size_t, POINTER(c_void_p)], c_int),
        "cft_program_free": ([V], None),
        "cft_program_get_info": ([V, POINTER(CftProgramInfo)], c_int),
        "cft_program_run": ([V, V, V, V, V, P, c_size_t, P, P], c_int),
        "cft_program_run_bank": ([V, V, c_size_t, V, V, V, V, P, c_size_t,
                                  P, P], c_int),
        "cft_program_run_ex": ([V, POINTER(CftRunArgs)], c_int),
        "cft_program_digest": ([V, V, c_size_t, V], c_int),
        "cftx_program_check": ([V, c_size_t, c_uint32, POINTER(CftProgramInfo)], c_int),
        "cftx_program_run_batch": ([V, POINTER(CftxBatchArgs)], c_int),
        "cftx_program_run_multi": ([V, POINTER(CftxBatchPart), c_size_t, V, P, P], c_int),
        "cftx_program_prepare": ([V, c_uint32, P], c_int),
        "cftx_multi_last_plan": ([V, POINTER(ctypes.c_uint64)], c_int),
        "cftx_batch_launches": ([V], ctypes.c_uint64),
        "cftx_batch_width": ([V], c_uint32),
        "cftx_server_open": ([V, c_int, c_size_t, POINTER(c_void_p)], c_int),
        "cftx_server_submit": ([V, POINTER(CftxJob), POINTER(ctypes.c_uint64)], c_int),
        "cftx_server_wait": ([V, ctypes.c_uint64, POINTER(CftxJobResult)], c_int),
        "cftx_server_clock": ([V], ctypes.c_double),
        "cftx_server_stats_get": ([V, POINTER(CftxServerStats)], c_int),
        "cftx_server_close": ([V], None),
        "cftx_server_set_batching": ([V, c_uint32], c_int),
        "cftx_server_set_slice": ([V, ctypes.c_double], c_int),
        "cftx_buffer_device_ptr": ([V, V, c_size_t], c_void_p),
        "cftx_reduce_noalloc": ([V, c_int, c_int, c_int, V, V, V, c_size_t, P, P], c_int),
        "cftx_program_wide": ([V, c_size_t, c_uint32, c_uint32, c_uint32, V, c_size_t,
                               POINTER(c_uint32)], c_int),
        "cft_lower_flags": ([V, c_uint32], None),
        "cft_raise_flags": ([V, c_uint32], None),
        "cft_test_flags": ([V, c_uint32], c_int),
        "cft_save_all_flags": ([V], c_uint32),
        "cft_restore_flags": ([V, c_uint32, c_uint32], None),
        "cft_test_saved_flags": ([c_uint32, c_uint32], c_int),
        "cftx_kernel_timer": ([V, c_int], c_int),
        "cftx_kernel_timer_read": ([V, POINTER(ctypes.c_double),
                                    POINTER(ctypes.c_uint64)], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def _audit(lib):
    abi = lib.cft_abi_version()
    if (abi >> 16) != ABI_MAJOR:
        raise RuntimeError(f"libcft reports ABI {abi >> 16}.{abi & 0xffff}; "
                           f"this binding is written against major {ABI_MAJOR}")
    wrong = []
    for name, value in sorted(globals().items()):
        if name.startswith("OP_"):
            want, got = name[3:].lower(), lib.cft_op_name(value)
        elif name in ("FP32", "FP64", "FP128", "FP256"):
            want, got = name.lower(), lib.cft_format_name(value)
        else:
            continue
        if got.decode() != want:
            wrong.append(f"{name}={value} is {got.decode()!r}")
    if wrong:
        raise RuntimeError("libcft disagrees with this binding's numbers: " +
                           ", ".join(wrong))


_LIB = None


def lib():
    """The bound, audited library (loaded once)."""
    global _LIB
    if _LIB is None:
        path = find_library()
        _LIB = _bind(ctypes.CDLL(path))
        _audit(_LIB)
    return _LIB


def library_path():
    return find_library()


def check(status, where):
    if status != OK:
        detail = lib().cft_last_error().decode(errors="replace")
        raise CftError(status, where, detail or lib().cft_strerror(status).decode())
