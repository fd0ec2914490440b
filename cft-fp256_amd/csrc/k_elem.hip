// Elementwise kernels: d[i] = op(a[i], b[i], c[i]) for the arithmetic group
// (FMA, ADD, SUB, MUL) and the sign/compare/integer groups, binary32 ..
// binary256.  The caller's arrays are dense little-endian elements (host/
// include/cft.h:1166-1209); each lane owns one element at a time, loads it
// with 16-byte vector loads (elem_io.hpp) and keeps it in VGPRs as 32-bit limbs.
//
// Replaces the reference's k_run / k_run256 / k_run128 (host/src/
// backend_hip.cpp:484-665).  Launch shape: a grid-stride loop, launched by
// the library with one workgroup for every 256 elements; each wavefront ORs
// its flags into one of the call's flag slots (kernels.h) with at most one
// atomic.
#include "elem_io.hpp"
#include "fx_arith.hpp"
#include "fx_misc.hpp"
#include "kernels.h"

using namespace fx;

namespace {

// One wavefront's flags into the call's word: the OR over its lanes by one
// ballot a flag bit, then one global atomic from its first active lane, and
// only for bits the word does not hold yet (after the first waves set them,
// most waves only read it).  No LDS, so these kernels can start on a CU whose
// LDS a persistent sequencer launch of another stream holds.
__device__ __forceinline__ void flush_flags(uint32_t fl, uint32_t *flags_out)
{
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

// Arithmetic group.  OP is FMA/ADD/SUB/MUL; operands NULL when not read.
template <int NL, int OP>
__global__ __launch_bounds__(256) void k_arith(const uint32_t *__restrict__ a,
                                               const uint32_t *__restrict__ b,
                                               const uint32_t *__restrict__ c,
                                               uint32_t *__restrict__ d, size_t n,
                                               uint32_t smask, int rnd,
                                               uint32_t *__restrict__ flags_out,
                                               uint32_t slot_mask)
{
    uint32_t fl = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t xa[NL], xb[NL], xc[NL], o[NL];
        load_el<NL>(a, (smask & 1u) ? 0 : i, xa);
        if (OP == OP_FMA || OP == OP_MUL)
            load_el<NL>(b, (smask & 2u) ? 0 : i, xb);
        else
#pragma unroll
            for (int k = 0; k < NL; k++) xb[k] = 0;
        if (OP != OP_MUL)
            load_el<NL>(c, (smask & 4u) ? 0 : i, xc);
        else
#pragma unroll
            for (int k = 0; k < NL; k++) xc[k] = 0;
        if constexpr (OP == OP_ADD || OP == OP_SUB)
            add_core<NL>(xa, xc, OP == OP_SUB, rnd, o, fl);   // a * 1 + c, one rounding
        else
            fma_op<NL, OP>(xa, xb, xc, rnd, o, fl);
        store_el<NL>(d, i, o);
    }
    flush_flags(fl, flags_out + (blockIdx.x & slot_mask) * CFTK_SLOT_WORDS);
}

// Every other elementwise opcode (sign, min/max, compare, select, integer,
// the seeds, and the unassigned ones): short, so one kernel switches on the
// uniform opcode.  NULL operands read as +0.
template <int NL>
__global__ __launch_bounds__(256) void k_misc(const uint32_t *__restrict__ a,
                                              const uint32_t *__restrict__ b,
                                              const uint32_t *__restrict__ c,
                                              uint32_t *__restrict__ d, size_t n,
                                              uint32_t smask, int op,
                                              uint32_t *__restrict__ flags_out,
                                              uint32_t slot_mask)
{
    uint32_t fl = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t xa[NL], xb[NL], xc[NL], o[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) xa[k] = xb[k] = xc[k] = 0;
        if (a) load_el<NL>(a, (smask & 1u) ? 0 : i, xa);
        if (b) load_el<NL>(b, (smask & 2u) ? 0 : i, xb);
        if (c) load_el<NL>(c, (smask & 4u) ? 0 : i, xc);
        misc_op<NL>(op, xa, xb, xc, o, fl);
        store_el<NL>(d, i, o);
    }
    flush_flags(fl, flags_out + (blockIdx.x & slot_mask) * CFTK_SLOT_WORDS);
}

__global__ void k_fill32(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// One wavefront an area (workgroup b, area b): lane s takes slot s's word
// and zeroes it, the wave ORs the words, lane 0 stores the OR to host
// memory.  The store is system scope, so it is in host memory when the
// stream's completion is seen.
__global__ __launch_bounds__(64) void k_flags_gather(uint32_t *slots, int nslots,
                                                     uint32_t *host_words)
{
    slots += (size_t)blockIdx.x * CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS;
    uint32_t *host_word = host_words + blockIdx.x;
    const int s = (int)threadIdx.x;
    uint32_t v = 0;
    if (s < nslots) {
        v = slots[s * CFTK_SLOT_WORDS];
        slots[s * CFTK_SLOT_WORDS] = 0u;
    }
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1)
        v |= __shfl_xor(v, k, 64);
    if (s == 0)
        __hip_atomic_store(host_word, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int NL>
hipError_t launch_arith(int op, const void *a, const void *b, const void *c, void *d, size_t n,
                        uint32_t smask, int rnd, uint32_t *flags, uint32_t sm, int grid,
                        hipStream_t st)
{
    const uint32_t *pa = (const uint32_t *)a, *pb = (const uint32_t *)b,
                   *pc = (const uint32_t *)c;
    uint32_t *pd = (uint32_t *)d;
    switch (op) {
    case OP_FMA:
        hipLaunchKernelGGL((k_arith<NL, OP_FMA>), dim3(grid), dim3(256), 0, st, pa, pb, pc, pd, n,
                           smask, rnd, flags, sm);
        break;
    case OP_ADD:
        hipLaunchKernelGGL((k_arith<NL, OP_ADD>), dim3(grid), dim3(256), 0, st, pa, pb, pc, pd, n,
                           smask, rnd, flags, sm);
        break;
    case OP_SUB:
        hipLaunchKernelGGL((k_arith<NL, OP_SUB>), dim3(grid), dim3(256), 0, st, pa, pb, pc, pd, n,
                           smask, rnd, flags, sm);
        break;
    default:
        hipLaunchKernelGGL((k_arith<NL, OP_MUL>), dim3(grid), dim3(256), 0, st, pa, pb, pc, pd, n,
                           smask, rnd, flags, sm);
        break;
    }
    return hipGetLastError();
}

template <int NL>
hipError_t launch_misc(int op, const void *a, const void *b, const void *c, void *d, size_t n,
                       uint32_t smask, uint32_t *flags, uint32_t sm, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_misc<NL>), dim3(grid), dim3(256), 0, st, (const uint32_t *)a,
                       (const uint32_t *)b, (const uint32_t *)c, (uint32_t *)d, n, smask, op,
                       flags, sm);
    return hipGetLastError();
}

} // namespace

extern "C" hipError_t cftk_elem(int fmt, int op, int rnd, const void *a, const void *b,
                                const void *c, void *d, size_t n, uint32_t smask,
                                uint32_t *flags, uint32_t slot_mask, int grid, hipStream_t st)
{
    const bool arith = op >= OP_FMA && op <= OP_MUL;
    const uint32_t sm = slot_mask;
    switch (fmt) {
    case 0: return arith ? launch_arith<1>(op, a, b, c, d, n, smask, rnd, flags, sm, grid, st)
                         : launch_misc<1>(op, a, b, c, d, n, smask, flags, sm, grid, st);
    case 1: return arith ? launch_arith<2>(op, a, b, c, d, n, smask, rnd, flags, sm, grid, st)
                         : launch_misc<2>(op, a, b, c, d, n, smask, flags, sm, grid, st);
    case 2: return arith ? launch_arith<4>(op, a, b, c, d, n, smask, rnd, flags, sm, grid, st)
                         : launch_misc<4>(op, a, b, c, d, n, smask, flags, sm, grid, st);
    default: return arith ? launch_arith<8>(op, a, b, c, d, n, smask, rnd, flags, sm, grid, st)
                          : launch_misc<8>(op, a, b, c, d, n, smask, flags, sm, grid, st);
    }
}

extern "C" hipError_t cftk_fill32(uint32_t *p, uint32_t v, size_t n, hipStream_t st)
{
    const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_fill32, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

extern "C" hipError_t cftk_flags_gather(uint32_t *slots, int nslots, uint32_t *host_words,
                                        int nareas, hipStream_t st)
{
    hipLaunchKernelGGL(k_flags_gather, dim3(nareas), dim3(64), 0, st, slots, nslots,
                       host_words);
    return hipGetLastError();
}
