// One element of a dense operand array (little-endian limbs, host/include/
// cft.h:1166-1209) to and from a lane's registers, for the elementwise and
// division/root kernels: 16-byte vector accesses, binary32 .. binary256.
#pragma once

#include "fx_hw.hpp"

namespace fx {

// A compiler barrier: memory accesses are not moved across it.
__device__ __forceinline__ void io_order()
{
    asm volatile("" ::: "memory");
}

// Element i of p into x.  The loads are issued in program order: a binary256
// element's low half, then its high half, then (at the barrier after the
// element) the next operand's.  Left to the scheduler, a change elsewhere in
// a kernel once reordered them as every operand's high half first, and the
// binary256 FMA kernel ran 4% slower at the same instruction count and
// occupancy; in this order it streams at the same-bytes copy rate.
template <int NL>
__device__ __forceinline__ void load_el(const uint32_t *__restrict__ p, size_t i,
                                        uint32_t (&x)[NL])
{
    if constexpr (NL == 8) {
        const uint4 *q = reinterpret_cast<const uint4 *>(p + i * 8);
        const uint4 v0 = q[0];
        io_order();
        const uint4 v1 = q[1];
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
        x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
    } else if constexpr (NL == 4) {
        const uint4 v0 = reinterpret_cast<const uint4 *>(p)[i];
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
    } else if constexpr (NL == 2) {
        const uint2 v0 = reinterpret_cast<const uint2 *>(p)[i];
        x[0] = v0.x; x[1] = v0.y;
    } else {
        x[0] = p[i];
    }
    io_order();
}

template <int NL>
__device__ __forceinline__ void store_el(uint32_t *__restrict__ p, size_t i,
                                         const uint32_t (&x)[NL])
{
    if constexpr (NL == 8) {
        uint4 *q = reinterpret_cast<uint4 *>(p + i * 8);
        q[0] = make_uint4(x[0], x[1], x[2], x[3]);
        q[1] = make_uint4(x[4], x[5], x[6], x[7]);
    } else if constexpr (NL == 4) {
        reinterpret_cast<uint4 *>(p)[i] = make_uint4(x[0], x[1], x[2], x[3]);
    } else if constexpr (NL == 2) {
        reinterpret_cast<uint2 *>(p)[i] = make_uint2(x[0], x[1]);
    } else {
        p[i] = x[0];
    }
}

} // namespace fx
