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
