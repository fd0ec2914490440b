"""The C-ABI library loads, exports every entry point include/*.h declares,
audits its opcode/format names, and applies the argument rules that need no
GPU.  No compute calls: this runs without a GPU."""

import ctypes
import os
import re
import subprocess

import pytest

from conftest import PKG, REPO

HEADERS = [os.path.join(REPO, "include", h) for h in ("cft.h", "cft_mi355x.h")]
LIB = os.path.join(PKG, "lib", "libcft.so")


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"CFT_API\s+[\w\s\*]*?\b(cftx?_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libcft.so is not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (cftx?_\w+)", out))
    syms = declared_symbols()
    assert len(syms) >= 50
    assert {"cftx_program_run_multi", "cftx_server_open", "cftx_program_run_batch"} <= set(syms)
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_binding_audit_and_names():
    from cftamd import _lib
    lib = _lib.lib()              # binds and audits
    assert lib.cft_abi_version() >> 16 == 0
    assert lib.cft_op_name(0) == b"fma" and lib.cft_op_name(15) == b"reserved"
    assert lib.cft_format_name(3) == b"fp256" and lib.cft_format_size(2) == 16
    assert lib.cft_strerror(1) == b"invalid argument"


def test_open_refusals_without_a_gpu():
    from cftamd import _lib
    lib = _lib.lib()
    h = ctypes.c_void_p()
    # no software backend in this library, by design
    assert lib.cft_open(None, 0, ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    assert b"no software backend" in lib.cft_last_error()
    assert lib.cft_open(b"hip:x", 0, ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_open(b"hip:0", 3, ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_open(b"tile.xclbin", 0, ctypes.byref(h)) == _lib.ERR_NO_DEVICE
    args = _lib.CftOpenArgs(ctypes.sizeof(_lib.CftOpenArgs), b"hip:0", 0, 3)
    assert lib.cft_open_ex(ctypes.byref(args), ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    args = _lib.CftOpenArgs(7, b"hip:0", 0, 0)
    assert lib.cft_open_ex(ctypes.byref(args), ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT


def test_argument_rules_before_any_device_work():
    from cftamd import _lib
    lib = _lib.lib()
    fl = ctypes.c_uint32(99)
    # NULL device
    assert lib.cft_run(None, 0, 3, 0, None, None, None, None, 1, ctypes.byref(fl),
                       None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_div(None, 3, 0, None, None, None, 1, None, None) == \
        _lib.ERR_INVALID_ARGUMENT
    assert lib.cft_run_ex(None, 0, 3, 0, None) == _lib.ERR_INVALID_ARGUMENT
    # the status-word operations tolerate a NULL device
    assert lib.cft_save_all_flags(None) == 0 and lib.cft_test_flags(None, 31) == 0
    assert lib.cft_test_saved_flags(16, 16) == 1 and lib.cft_test_saved_flags(8, 16) == 0


@pytest.mark.skipif(not os.path.exists("/root/reference/host/include/cft.h"),
                    reason="the reference header is only in the build container")
def test_struct_layouts_match_reference(tmp_path):
    """sizeof/offsetof of every ABI struct against the reference header."""
    probe = r'''
#include <stdio.h>
#include <stddef.h>
#include "cft.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t,f) printf(#t "." #f " %zu\n", offsetof(t,f));
int main(void){
 S(cft_caps) O(cft_caps,backend) O(cft_caps,max_deposits) O(cft_caps,buffers_resident)
 S(cft_open_args) O(cft_open_args,scratch_depth)
 S(cft_buffer_info) O(cft_buffer_info,staged_why)
 S(cft_program_info) O(cft_program_info,scratch_used)
 S(cft_run_args) O(cft_run_args,bank) O(cft_run_args,idx_a) O(cft_run_args,lane_mask)
 O(cft_run_args,lane_flags_bytes)
 S(cft_elem_args) O(cft_elem_args,scalar_mask) O(cft_elem_args,idx_c_src)
 printf("%d %d %d %d\n", (int)CFT_MAXALL, (int)CFT_ERR_INTERNAL, (int)CFT_RMM,
        (int)CFT_SEQ_FEAT_FLAG_CONTROL);
 return 0; }
'''
    src = tmp_path / "probe.c"
    src.write_text(probe)
    outs = []
    for inc in (os.path.join(REPO, "include"), "/root/reference/host/include"):
        exe = tmp_path / ("probe_" + str(len(outs)))
        subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True).stdout)
    assert outs[0] == outs[1]


# Every entry point the reference's Python binding looks up when it loads a
# library (bindings/python/cftmpfr/_lib.py:181-376, `_bind`), read from that
# file as text; the binding fails at load on the first one missing.
BIND_NAMES = {
    # _lib.py:190-231: names, devices, the batch calls
    "core": ["cft_abi_version", "cft_strerror", "cft_last_error", "cft_format_size",
             "cft_format_name", "cft_op_name", "cft_open", "cft_close", "cft_get_caps",
             "cft_supports", "cft_run", "cft_reduce", "cft_div", "cft_sqrt"],
    # _lib.py:233-247: one-operand transcendentals
    "unary": ["cft_exp", "cft_expm1", "cft_exp2", "cft_log", "cft_log1p", "cft_log2",
              "cft_log10", "cft_sinpi", "cft_cospi", "cft_tanpi", "cft_asin", "cft_acos",
              "cft_atan", "cft_asinpi", "cft_acospi", "cft_atanpi", "cft_sin", "cft_cos",
              "cft_tan", "cft_sinh", "cft_cosh", "cft_tanh", "cft_asinh", "cft_acosh",
              "cft_atanh", "cft_exp2m1", "cft_exp10", "cft_exp10m1", "cft_log2p1",
              "cft_log10p1", "cft_rsqrt"],
    # _lib.py:248-253
    "binary": ["cft_pow", "cft_hypot", "cft_atan2", "cft_atan2pi", "cft_powr"],
    # _lib.py:258-263
    "intexp": ["cft_pown", "cft_compound", "cft_rootn"],
    # _lib.py:269-275
    "augmented": ["cft_augmented_add", "cft_augmented_sub", "cft_augmented_mul"],
    # _lib.py:281-291
    "scaled": ["cft_scaled_prod", "cft_scaled_prod_sum", "cft_scaled_prod_diff"],
    # _lib.py:297-317
    "formatof": ["cft_formatof_add", "cft_formatof_sub", "cft_formatof_mul",
                 "cft_formatof_div", "cft_formatof_sqrt", "cft_formatof_fma"],
    # _lib.py:322-338
    "chars": ["cft_format_decimal_digits", "cft_from_decimal_char", "cft_from_hex_char",
              "cft_to_decimal_char", "cft_to_hex_char"],
    "payload": ["cft_get_payload", "cft_set_payload", "cft_set_payload_signaling"],
    # _lib.py:340-345 (MINMAX_MAG, _lib.py:99)
    "mag": ["cft_min_mag", "cft_max_mag", "cft_minnum_mag", "cft_maxnum_mag"],
    # _lib.py:351-362
    "flags": ["cft_lower_flags", "cft_raise_flags", "cft_test_flags", "cft_save_all_flags",
              "cft_restore_flags", "cft_test_saved_flags"],
    # _lib.py:366-370
    "conformance": ["cft_is754version1985", "cft_is754version2008", "cft_is754version2019"],
}


def test_reference_binding_finds_every_name():
    """What `_bind` does first - getattr of every name - succeeds, and each
    name is an exported function of libcft.so (no reference code runs)."""
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (cft_\w+)", out))
    names = [n for group in BIND_NAMES.values() for n in group]
    assert len(names) == len(set(names)) == 86
    lib = ctypes.CDLL(LIB)
    for n in names:
        assert n in exported, n
        getattr(lib, n)


def test_off_path_entry_points_refuse_by_name():
    """The entry points off this path return CFT_ERR_UNSUPPORTED with a
    sentence naming themselves, and write nothing; the two formulas are
    answered."""
    from cftamd import _lib
    lib = _lib.lib()
    V = ctypes.c_void_p
    d = (ctypes.c_uint8 * 64)(*([0xA5] * 64))
    fl = ctypes.c_uint32(0xDEAD)
    sig = {"unary": 7, "binary": 8, "intexp": 8, "augmented": 8, "scaled": 9,
           "payload": 5, "mag": 7}
    calls = 0
    for group, argc in sig.items():
        for name in BIND_NAMES[group]:
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            # the device NULL, d an output buffer, the rest NULL or 0
            args = [None] * (8 if name == "cft_scaled_prod" else argc)
            args[3 if group in ("augmented", "payload", "mag") else -3] = ctypes.cast(d, V)
            st = fn(*args)
            assert st == _lib.ERR_UNSUPPORTED, name
            assert name.encode() in lib.cft_last_error(), name
            calls += 1
    for name in BIND_NAMES["formatof"]:
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        argc = {"cft_formatof_sqrt": 9, "cft_formatof_fma": 11}.get(name, 10)
        assert fn(*([None] * argc)) == _lib.ERR_UNSUPPORTED, name
        calls += 1
    for name, argc in (("cft_from_decimal_char", 8), ("cft_from_hex_char", 8),
                       ("cft_to_decimal_char", 9), ("cft_to_hex_char", 6)):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        assert fn(*([None] * argc)) == _lib.ERR_UNSUPPORTED, name
        calls += 1
    assert calls == 62
    assert bytes(d) == b"\xa5" * 64 and fl.value == 0xDEAD
    lib.cft_format_decimal_digits.restype = ctypes.c_size_t
    assert [lib.cft_format_decimal_digits(f) for f in (0, 1, 2, 3, 4, -1)] == \
        [9, 17, 36, 73, 0, 0]
    for n in BIND_NAMES["conformance"]:
        assert getattr(lib, n)() == 0


def test_product_library_has_no_plant_code():
    """The fail-launch plant exists only in lib/libcft_plants.so (host code
    built -DCFT_PLANTS); libcft.so never reads CFT_PLANT."""
    assert b"CFT_PLANT" not in open(LIB, "rb").read()
    plants = os.path.join(PKG, "lib", "libcft_plants.so")
    if os.path.exists(plants):
        assert b"CFT_PLANT" in open(plants, "rb").read()
