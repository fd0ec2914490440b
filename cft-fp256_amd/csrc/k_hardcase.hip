// The hard-case generator (hardcase.hpp) over a chunk of cases, on the GPU.
// Built into lib/libcftprobe.so, a test-support library beside libcft.so
// (which does not contain it).
#include <hip/hip_runtime.h>
#include <stdint.h>

// The generator inlines the arithmetic's limb shifts many times over; where
// the optimizer declines to unroll one of them the loop stays, correct and
// slower, and the warning says nothing about this file's results.
#pragma clang diagnostic ignored "-Wpass-failed"

#include "hardcase.hpp"

using namespace hc;

namespace {

template <int NL>
__global__ __launch_bounds__(256) void k_hardcase(int op, uint64_t seed, uint64_t base, size_t n,
                                                  uint32_t *__restrict__ A,
                                                  uint32_t *__restrict__ B,
                                                  uint32_t *__restrict__ C,
                                                  uint8_t *__restrict__ fam_out)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t a[NL], b[NL], c[NL];
        const int fam = case_at<NL>(seed, base + i, op, a, b, c);
#pragma unroll
        for (int k = 0; k < NL; k++) {
            A[i * NL + k] = a[k];
            if (B)
                B[i * NL + k] = b[k];
            if (C)
                C[i * NL + k] = c[k];
        }
        if (fam_out)
            fam_out[i] = (uint8_t)fam;
    }
}

template <int NL>
hipError_t launch(int op, uint64_t seed, uint64_t base, size_t n, void *a, void *b, void *c,
                  uint8_t *fam)
{
    const size_t blocks = (n + 255) / 256;
    const int grid = (int)(blocks < 16384 ? blocks : 16384);
    hipLaunchKernelGGL((k_hardcase<NL>), dim3(grid), dim3(256), 0, 0, op, seed, base, n,
                       (uint32_t *)a, (uint32_t *)b, (uint32_t *)c, fam);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : hipDeviceSynchronize();
}

} // namespace

// n cases g = base .. base+n-1 of `seed` for op (0 FMA, 1 ADD, 2 SUB, 3 MUL,
// 256 DIV, 257 SQRT) in format fmt (0 binary32 .. 3 binary256), into device
// arrays a (always), b (FMA, MUL, DIV) and c (FMA, ADD, SUB) of n elements,
// and each case's family into fam (n bytes) where fam is not NULL.  A
// pointer the op does not read may be NULL.  Synchronous; returns the HIP
// error code (0 on success), or -1 for an argument it refuses.
extern "C" __attribute__((visibility("default"))) int
cftp_hardcases(int fmt, int op, uint64_t seed, uint64_t base, size_t n, void *a, void *b,
               void *c, uint8_t *fam)
{
    const bool needb = op_reads_b(op), needc = op_reads_c(op);
    if (fmt < 0 || fmt > 3 || !op_known(op))
        return -1;
    if (n == 0)
        return 0;
    if (!a || (needb && !b) || (needc && !c))
        return -1;
    if (!needb)
        b = nullptr;
    if (!needc)
        c = nullptr;
    switch (fmt) {
    case 0: return (int)launch<1>(op, seed, base, n, a, b, c, fam);
    case 1: return (int)launch<2>(op, seed, base, n, a, b, c, fam);
    case 2: return (int)launch<4>(op, seed, base, n, a, b, c, fam);
    default: return (int)launch<8>(op, seed, base, n, a, b, c, fam);
    }
}
