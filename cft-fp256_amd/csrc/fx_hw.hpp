// The gfx950 instructions the limb arithmetic is written around, two binary64
// estimates, and the HIP runtime header.  Included as <fx_hw.hpp> so that a portable CPU build of
// the same arithmetic (the test suite's) can supply its own definitions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fx {

// ((hi:lo) >> (s & 31)) [31:0]: v_alignbit_b32.  Written as the instruction
// itself: hipcc 7.2's gfx950 backend has been seen to lower a funnel shift of
// limb x by k to v_alignbit on the register holding x << k when x << k is
// also live (the IR was right, the ISA was not), so every multi-limb shift in
// this library goes through this opaque form.
__device__ __forceinline__ uint32_t alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    uint32_t r;
    asm("v_alignbit_b32 %0, %1, %2, %3" : "=v"(r) : "v"(hi), "v"(lo), "v"(s));
    return r;
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

// The same for the first product of a column: hi = the carry out, not hi +
// it, so a column needs no instruction to clear hi first.
__device__ __forceinline__ void mac_first(uint64_t &acc, uint32_t &hi, uint32_t a, uint32_t b)
{
    uint64_t cc;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc), "=s"(cc) : "v"(a), "v"(b));
    asm("v_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(hi) : "s"(cc));
}

// x + y + cin (cin 0 or 1), carry out: one v_add_co / v_addc in a chain.
__device__ __forceinline__ uint32_t addc(uint32_t x, uint32_t y, uint32_t cin, uint32_t *cout)
{
    return __builtin_addc(x, y, cin, cout);
}

// x - y - bin (bin 0 or 1), borrow out: one v_sub_co / v_subb in a chain.
__device__ __forceinline__ uint32_t subc(uint32_t x, uint32_t y, uint32_t bin, uint32_t *bout)
{
    return __builtin_subc(x, y, bin, bout);
}

// 1/x and sqrt(x), correctly rounded, for x between 2^62 and 2^66 (the digit
// estimates' operands): the compiler's own sequences for 1.0 / x and sqrt(x)
// - a hardware estimate, Newton steps, a final fused correction - without
// their range scaling and special-case fix-ups, which leave a finite x of
// this size unchanged.  The same bits as the division and the root.
__device__ __forceinline__ double recip_mid(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(e, y, y);
}

__device__ __forceinline__ double sqrt_mid(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, s, 0.5);
    s = __builtin_fma(s, r, s);
    double d = __builtin_fma(-s, s, x);
    h = __builtin_fma(h, r, h);
    s = __builtin_fma(d, h, s);
    d = __builtin_fma(-s, s, x);
    return __builtin_fma(d, h, s);
}

} // namespace fx
