// Launchers for the HIP kernels, called by the host side of the C ABI.
// Every pointer is a device pointer; every launch goes on `st`.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

extern "C" {

// d[i] = op(a[i], b[i], c[i]) over n elements of format fmt (0..3).  An
// operand the opcode does not read may be NULL; bit r of smask makes operand
// r a single element.  The call's flags are ORed into *flags.
hipError_t cftk_elem(int fmt, int op, int rnd, const void *a, const void *b, const void *c,
                     void *d, size_t n, uint32_t smask, uint32_t *flags, int grid,
                     hipStream_t st);

// Correctly-rounded a/b (op 0) or sqrt(a) (op 1).
hipError_t cftk_divsqrt(int fmt, int which, int rnd, const void *a, const void *b, void *d,
                        size_t n, uint32_t smask, uint32_t *flags, int grid, hipStream_t st);

// p[0..n) = v
hipError_t cftk_fill32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);

}
