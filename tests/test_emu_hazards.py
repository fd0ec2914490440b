"""Hazard builds of the device limb arithmetic, on the CPU: tests/emu's
build of cft-fp256_amd/csrc/fx_*.hpp checked against the oracle element by
element (tests/emu/emu_check.cpp) when compiled at -O0 (no optimisation to
hide an uninitialised read or rely on a folded constant) and under
AddressSanitizer + UndefinedBehaviorSanitizer (out-of-bounds limb indexing,
shifts by the word size or more, signed overflow).  Host code only; these
sanitizers never touch GPU code."""

import os
import subprocess

import pytest

from conftest import EMU_DIR, PKG, REPO

ORACLE_C = [os.path.join(REPO, "oracle", f) for f in ("sf_oracle.c", "seq_oracle.c",
                                                      "rd_oracle.c")]


def build(tmp_path, tag, flags):
    objs = []
    for src in ORACLE_C:
        obj = tmp_path / f"{tag}_{os.path.basename(src)}.o"
        subprocess.run(["gcc", "-std=c99", "-c", src, "-o", str(obj)] + flags, check=True)
        objs.append(str(obj))
    exe = tmp_path / f"emu_check_{tag}"
    subprocess.run(["g++", "-std=c++17", "-Wno-unknown-pragmas", "-I", EMU_DIR,
                    "-I", os.path.join(PKG, "csrc"), os.path.join(EMU_DIR, "emu_check.cpp"),
                    "-o", str(exe), "-pthread"] + objs + flags, check=True)
    return exe


@pytest.mark.parametrize("tag,flags,cases", [
    ("O0", ["-O0", "-g"], 600),
    ("asan_ubsan", ["-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"], 300),
])
def test_emu_against_oracle_in_hazard_build(tmp_path, tag, flags, cases):
    exe = build(tmp_path, tag, flags)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([str(exe), str(cases), "12345"], capture_output=True, text=True,
                       timeout=600, env=env)
    assert p.returncode == 0, (p.stdout[-1000:], p.stderr[-3000:])
    assert "0 mismatches" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
