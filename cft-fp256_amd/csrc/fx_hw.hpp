// The two gfx950 instructions the limb arithmetic is written around, and the
// HIP runtime header.  Included as <fx_hw.hpp> so that a portable CPU build of
// the same arithmetic (the test suite's) can supply its own definitions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fx {

// ((hi:lo) >> (s & 31)) [31:0]: v_alignbit_b32
__device__ __forceinline__ uint32_t alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    return __builtin_amdgcn_alignbit(hi, lo, s);
}

// acc64 += a * b, the carry out of bit 63 added into hi: one v_mad_u64_u32
// with its carry-out and one v_addc.  The compiler's own lowering of this
// 96-bit accumulate costs four to five instructions.
__device__ __forceinline__ void mac(uint64_t &acc, uint32_t &hi, uint32_t a, uint32_t b)
{
    uint64_t cc;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc), "=s"(cc) : "v"(a), "v"(b));
    asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(hi), "=s"(cc) : "s"(cc));
}

} // namespace fx
