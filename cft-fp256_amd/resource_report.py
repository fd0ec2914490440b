#!/usr/bin/env python3
"""Per-kernel resources of every HIP kernel in libcft.so (and the test-support
kernels of lib/libcftprobe.so beside it), from the compiler itself: each kernel source is compiled
for gfx950 with the library's own device flags and
-Rpass-analysis=kernel-resource-usage, and the remarks - VGPRs, AGPRs,
SGPRs, scratch (spill) bytes a lane, static LDS, the compiler's occupancy
bound in waves a SIMD - are printed as a markdown table.

    python3 cft-fp256_amd/resource_report.py > profiles/kernel_resources.md

Dynamic LDS (the sequencer kernels size their register files and banks at
launch) is not in the static figure; the launchers' own LDS formulas
(cftk_seq_lds, cftk_seq_wide_lds, cftk_seq_multi_lds in csrc/k_seq.hip) give
it per program."""

import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["k_elem.hip", "k_divsqrt.hip", "k_reduce.hip", "k_seq.hip", "cft_noalloc.hip",
           "k_probe.hip", "k_chain.hip", "k_hardcase.hip", "k_lp.hip"]
FLAGS = ["--offload-arch=gfx950", "-Icsrc", "-O3", "-std=c++17", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "--offload-device-only", "-c", "-o", os.devnull,
         "-Rpass-analysis=kernel-resource-usage"]
FIELDS = [("VGPRs", "VGPRs"), ("AGPRs", "AGPRs"), ("TotalSGPRs", "SGPRs"),
          ("ScratchSize [bytes/lane]", "scratch B/lane"), ("VGPRs Spill", "VGPR spill"),
          ("SGPRs Spill", "SGPR spill"), ("LDS Size [bytes/block]", "static LDS B"),
          ("Occupancy [waves/SIMD]", "waves/SIMD")]


def demangle(names):
    tool = shutil.which("c++filt") or shutil.which("llvm-cxxfilt")
    if not tool:
        return names
    out = subprocess.run([tool], input="\n".join(names), capture_output=True, text=True)
    got = out.stdout.splitlines()
    return got if len(got) == len(names) else names


def short(name):
    name = name.replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*\)$", "", name)


def main():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    rows = []
    for src in SOURCES:
        p = subprocess.run([hipcc] + FLAGS + [os.path.join("csrc", src)], cwd=HERE,
                           capture_output=True, text=True)
        if p.returncode != 0:
            sys.exit(f"compiling {src} failed:\n{p.stderr[-2000:]}")
        cur = None
        for line in p.stderr.splitlines():
            m = re.search(r"remark: Function Name: (\S+)", line)
            if m:
                cur = {"src": src, "name": m.group(1)}
                rows.append(cur)
                continue
            m = re.search(r"remark:\s+([^:]+): (\S+) \[-Rpass-analysis", line)
            if m and cur is not None:
                cur[m.group(1).strip()] = m.group(2)
    names = demangle([r["name"] for r in rows])
    print("# Per-kernel resources (gfx950, from the compiler)\n")
    print("Every kernel of libcft.so and lib/libcftprobe.so, compiled with the library's "
          "device flags and\n`-Rpass-analysis=kernel-resource-usage` (hipcc, ROCm 7.2). "
          "Command, from the repository root:\n")
    print("    python3 cft-fp256_amd/resource_report.py > profiles/kernel_resources.md\n")
    spilled = [short(n) for r, n in zip(rows, names)
               if r.get("ScratchSize [bytes/lane]", "0") not in ("0", "?")]
    print("waves/SIMD is the compiler's bound from registers and static LDS at the "
          "launch bounds; the\nsequencer kernels also take dynamic LDS sized at launch "
          "(cftk_seq_lds, cftk_seq_wide_lds,\ncftk_seq_multi_lds), which bounds them "
          "lower per program.  " +
          ("No kernel spills (scratch 0)." if not spilled else
           "Kernels with scratch (spilled registers): " + ", ".join(f"`{k}`" for k in spilled) +
           ".") + "\n")
    print("| source | kernel | " + " | ".join(h for _k, h in FIELDS) + " |")
    print("|---|---|" + "---|" * len(FIELDS))
    for r, n in zip(rows, names):
        print(f"| {r['src']} | `{short(n)}` | " +
              " | ".join(r.get(k, "?") for k, _h in FIELDS) + " |")


if __name__ == "__main__":
    main()
