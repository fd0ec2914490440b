// Correctly-rounded division and square root kernels, binary32 .. binary256,
// one lane per element.  Replace the reference's composed routes (host/src/
// divsqrt.c: ~25-30 elementwise passes, or a Newton program) with one pass:
// the contract is "correctly rounded", so the bits and flags are the same.
#include "elem_io.hpp"
#include "fx_divsqrt.hpp"
#include "kernels.h"

using namespace fx;

namespace {

template <int NL, int WHICH>
__global__ __launch_bounds__(256) void k_divsqrt(const uint32_t *__restrict__ a,
                                                 const uint32_t *__restrict__ b,
                                                 uint32_t *__restrict__ d, size_t n, int rnd,
                                                 uint32_t *__restrict__ flags_out,
                                                 uint32_t slot_mask)
{
    uint32_t fl = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t xa[NL], xb[NL], o[NL];
        load_el<NL>(a, i, xa);
        if (WHICH == 0) {
            load_el<NL>(b, i, xb);
            div_op<NL>(xa, xb, rnd, o, fl);
        } else {
            sqrt_op<NL>(xa, rnd, o, fl);
        }
        store_el<NL>(d, i, o);
    }
    // the wavefront's OR of its lanes' flags, one atomic into its flag slot
    // for bits not yet set (no LDS: see flush_flags in k_elem.hip)
    flags_out += (blockIdx.x & slot_mask) * CFTK_SLOT_WORDS;
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 5; b++)
        if (__ballot((fl >> b) & 1u))
            w |= 1u << b;
    const uint64_t live = __ballot(1);
    if (w && __lane_id() == (unsigned)__builtin_ctzll(live)) {
        const uint32_t have = __hip_atomic_load(flags_out, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
        if (w & ~have)
            atomicOr(flags_out, w);
    }
}

template <int NL>
hipError_t launch(int which, int rnd, const void *a, const void *b, void *d, size_t n,
                  uint32_t *flags, uint32_t sm, int grid, hipStream_t st)
{
    if (which == 0)
        hipLaunchKernelGGL((k_divsqrt<NL, 0>), dim3(grid), dim3(256), 0, st,
                           (const uint32_t *)a, (const uint32_t *)b, (uint32_t *)d, n, rnd,
                           flags, sm);
    else
        hipLaunchKernelGGL((k_divsqrt<NL, 1>), dim3(grid), dim3(256), 0, st,
                           (const uint32_t *)a, (const uint32_t *)b, (uint32_t *)d, n, rnd,
                           flags, sm);
    return hipGetLastError();
}

} // namespace

extern "C" hipError_t cftk_divsqrt(int fmt, int which, int rnd, const void *a, const void *b,
                                   void *d, size_t n, uint32_t smask, uint32_t *flags,
                                   uint32_t slot_mask, int grid, hipStream_t st)
{
    (void)smask;
    switch (fmt) {
    case 0: return launch<1>(which, rnd, a, b, d, n, flags, slot_mask, grid, st);
    case 1: return launch<2>(which, rnd, a, b, d, n, flags, slot_mask, grid, st);
    case 2: return launch<4>(which, rnd, a, b, d, n, flags, slot_mask, grid, st);
    default: return launch<8>(which, rnd, a, b, d, n, flags, slot_mask, grid, st);
    }
}
