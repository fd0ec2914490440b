// Latency probe: one wavefront runs a dependent chain of binary256
// operations in registers (each result the next operation's operand), so the
// time an operation takes is its latency alone, with no memory traffic and no
// interpreter around it.  Built into lib/libcftprobe.so, a measurement aid
// beside libcft.so (which does not contain it).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fx_arith.hpp"

using namespace fx;

namespace {

// KIND 0: x <- x*a + c (fma_core); 1: x <- x + c (add_core, the two-term
// path); 2: x <- x*a + (+0) (a MUL steered to the FMA path)
template <int KIND>
__global__ __launch_bounds__(64) void k_chain(uint32_t *io, int iters)
{
    uint32_t x[8], a[8], c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        x[k] = io[k];
        a[k] = io[8 + k];
        c[k] = KIND == 2 ? 0u : io[16 + k];
    }
    uint32_t fl = 0;
    for (int it = 0; it < iters; it++) {
        uint32_t o[8];
        if constexpr (KIND == 1)
            add_core<8>(x, c, false, RNE, o, fl);
        else
            fma_core<8>(OP_FMA, x, a, c, RNE, o, fl);
#pragma unroll
        for (int k = 0; k < 8; k++)
            x[k] = o[k];
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 8; k++)
            io[24 + k] = x[k];
        io[32] = fl;
    }
}

template <int KIND>
hipError_t launch(uint32_t *io, int iters, hipStream_t st)
{
    hipLaunchKernelGGL((k_chain<KIND>), dim3(1), dim3(64), 0, st, io, iters);
    return hipGetLastError();
}

} // namespace

// One wavefront, `iters` dependent operations of `kind` (above) from the 24
// words in (x, a, c: binary256 encodings, limbs little end first); out gets
// the final x (8 words) and the flags (1 word); *us_per_op the kernel time
// (HIP events, after one untimed launch of 16) over iters.  Returns 0 or the
// HIP error code.
extern "C" __attribute__((visibility("default"))) int cftp_chain_latency(int kind, int iters,
                                                                        const uint32_t *in,
                                                                        uint32_t *out,
                                                                        double *us_per_op)
{
    if (kind < 0 || kind > 2 || iters < 1 || !in || !out || !us_per_op)
        return (int)hipErrorInvalidValue;
    uint32_t *io = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&io, 33 * 4);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto run = [&](int n) -> hipError_t {
        hipError_t r = hipMemcpyAsync(io, in, 24 * 4, hipMemcpyHostToDevice, st);
        if (r != hipSuccess)
            return r;
        switch (kind) {
        case 0: return launch<0>(io, n, st);
        case 1: return launch<1>(io, n, st);
        default: return launch<2>(io, n, st);
        }
    };
    if (e == hipSuccess) e = run(16);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMemcpyAsync(io, in, 24 * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) {
        switch (kind) {
        case 0: e = launch<0>(io, iters, st); break;
        case 1: e = launch<1>(io, iters, st); break;
        default: e = launch<2>(io, iters, st); break;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(out, io + 24, 9 * 4, hipMemcpyDeviceToHost);
    *us_per_op = (double)t * 1e3 / iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (io) (void)hipFree(io);
    return (int)e;
}

namespace {

// One wavefront, alone on its SIMD: `iters` rounds of 32 instructions of one
// kind spread over CH independent chains (CH 1: each instruction waits for
// the one before; CH 8: eight in flight), timed in shader clocks by the wave
// itself (s_memtime).  Cycles an instruction, for the two limits of a lone
// wavefront: issue (many chains) and dependent latency (one chain).
template <int KIND, int CH>
__global__ __launch_bounds__(64) void k_lone(uint32_t *out, unsigned long long *cyc,
                                             uint32_t seed, int iters)
{
    const uint32_t t = threadIdx.x;
    const uint32_t a = seed ^ t, b = seed * 2654435761u + t;
    uint64_t acc[CH];
    uint32_t v[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        acc[k] = (uint64_t)(t + k) << 7;
        v[k] = t * (k + 3);
    }
    const unsigned long long t0 = clock64();
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const int k = q % CH;
            if constexpr (KIND == 1) {
                uint64_t cc;
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0"
                             : "+v"(acc[k]), "=s"(cc)
                             : "v"(a), "v"(b));
            } else if constexpr (KIND == 2) {
                // a lane's value from the lane one up, through the LDS crossbar
                v[k] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(((t + 1) & 63) << 2),
                                                              (int)v[k]);
            } else if constexpr (KIND == 3) {
                // the same through the data-parallel move (row_shr:1)
                v[k] = (uint32_t)__builtin_amdgcn_update_dpp((int)v[k], (int)v[k], 0x111, 0xF,
                                                             0xF, false);
            } else {
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(v[k]) : "v"(a));
            }
        }
    }
    const unsigned long long t1 = clock64();
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < CH; k++)
        s ^= (uint32_t)acc[k] ^ (uint32_t)(acc[k] >> 32) ^ v[k];
    out[t] = s;
    if (t == 0)
        *cyc = t1 - t0;
}

template <int KIND>
hipError_t launch_lone(int chains, uint32_t *out, unsigned long long *cyc, int iters)
{
    switch (chains) {
    case 1: hipLaunchKernelGGL((k_lone<KIND, 1>), dim3(1), dim3(64), 0, 0, out, cyc, 7u, iters); break;
    case 2: hipLaunchKernelGGL((k_lone<KIND, 2>), dim3(1), dim3(64), 0, 0, out, cyc, 7u, iters); break;
    case 4: hipLaunchKernelGGL((k_lone<KIND, 4>), dim3(1), dim3(64), 0, 0, out, cyc, 7u, iters); break;
    default: hipLaunchKernelGGL((k_lone<KIND, 8>), dim3(1), dim3(64), 0, 0, out, cyc, 7u, iters); break;
    }
    return hipGetLastError();
}

} // namespace

// Shader clocks an instruction for one lone wavefront: kind 0 v_add_u32, 1
// v_mad_u64_u32, 2 ds_bpermute_b32, 3 v_mov_b32 with a DPP row shift; chains
// 1, 2, 4 or 8 independent dependence chains.
// Returns 0 or the HIP error code.
extern "C" __attribute__((visibility("default"))) int cftp_lone_issue(int kind, int chains,
                                                                     int iters,
                                                                     double *clocks_per_insn)
{
    if (kind < 0 || kind > 3 || !(chains == 1 || chains == 2 || chains == 4 || chains == 8) ||
        iters < 1 || !clocks_per_insn)
        return (int)hipErrorInvalidValue;
    uint32_t *out = nullptr;
    unsigned long long *cyc = nullptr;
    hipError_t e = hipMalloc(&out, 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&cyc, sizeof *cyc);
    for (int rep = 0; e == hipSuccess && rep < 2; rep++) {     // the first warms the code
        switch (kind) {
        case 0: e = launch_lone<0>(chains, out, cyc, iters); break;
        case 1: e = launch_lone<1>(chains, out, cyc, iters); break;
        case 2: e = launch_lone<2>(chains, out, cyc, iters); break;
        default: e = launch_lone<3>(chains, out, cyc, iters); break;
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    unsigned long long c = 0;
    if (e == hipSuccess) e = hipMemcpy(&c, cyc, sizeof c, hipMemcpyDeviceToHost);
    *clocks_per_insn = (double)c / ((double)iters * 32.0);
    if (cyc) (void)hipFree(cyc);
    if (out) (void)hipFree(out);
    return (int)e;
}
