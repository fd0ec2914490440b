// Issue-rate probes for the roofline: how many wave instructions a second
// gfx950 issues of the integer operations the arithmetic is built from.
// Built into lib/libcftprobe.so, a measurement aid beside libcft.so (which
// does not contain it).
//
// Each lane runs `iters` rounds of 8 independent chains of one instruction
// (so issue, not latency, is the limit once a SIMD holds a few waves); the
// instructions are written as inline asm, so exactly the instruction named
// is issued, once per chain and round.  The chains' values are stored at
// the end so nothing is dead.  A launch issues blocks * 4 waves * iters * 8
// wave instructions of the probed kind (plus a few of loop control).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

enum { P_ADD = 0, P_MAD64 = 1, P_MULLO = 2, P_MULHI = 3, P_ADDC = 4 };

template <int KIND>
__global__ __launch_bounds__(256) void k_issue(uint32_t *out, uint32_t seed, int iters)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a = seed ^ t, b = seed * 2654435761u + t;
    uint64_t acc[8], cc[8];
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        acc[k] = (uint64_t)(t + k) << 7;
        v[k] = t * (k + 3);
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if constexpr (KIND == P_MAD64) {
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0"
                             : "+v"(acc[k]), "=s"(cc[k])
                             : "v"(a), "v"(b));
            } else if constexpr (KIND == P_ADD) {
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(v[k]) : "v"(a));
            } else if constexpr (KIND == P_MULLO) {
                asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(v[k]) : "v"(b));
            } else if constexpr (KIND == P_MULHI) {
                asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(v[k]) : "v"(b));
            } else {
                asm volatile("v_add_co_u32 %0, vcc, %0, %1\n\t"
                             "v_addc_co_u32 %0, vcc, %0, %1, vcc"
                             : "+v"(v[k])
                             : "v"(a)
                             : "vcc");
            }
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        s ^= (uint32_t)acc[k] ^ (uint32_t)(acc[k] >> 32) ^ v[k];
    out[t] = s;
}

template <int KIND>
hipError_t launch(uint32_t *out, int blocks, int iters, hipStream_t st)
{
    hipLaunchKernelGGL((k_issue<KIND>), dim3(blocks), dim3(256), 0, st, out, 12345u, iters);
    return hipGetLastError();
}

} // namespace

// Times `reps` launches of the probe `kind` (0 v_add_u32, 1 v_mad_u64_u32,
// 2 v_mul_lo_u32, 3 v_mul_hi_u32, 4 a v_add_co_u32 + v_addc_co_u32 pair)
// over `blocks` workgroups of 256 lanes after one untimed launch; *ms gets
// the mean kernel time (HIP events on the probe's own stream) and
// *wave_insns the probed wave instructions one launch issues.  Returns 0 or
// the HIP error code.
extern "C" __attribute__((visibility("default"))) int cftp_issue_rate(int kind, int blocks,
                                                                      int iters, int reps,
                                                                      double *ms,
                                                                      double *wave_insns)
{
    if (kind < 0 || kind > 4 || blocks < 1 || iters < 1 || reps < 1 || !ms || !wave_insns)
        return (int)hipErrorInvalidValue;
    uint32_t *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)blocks * 256 * 4);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto run = [&]() -> hipError_t {
        switch (kind) {
        case P_ADD: return launch<P_ADD>(out, blocks, iters, st);
        case P_MAD64: return launch<P_MAD64>(out, blocks, iters, st);
        case P_MULLO: return launch<P_MULLO>(out, blocks, iters, st);
        case P_MULHI: return launch<P_MULHI>(out, blocks, iters, st);
        default: return launch<P_ADDC>(out, blocks, iters, st);
        }
    };
    if (e == hipSuccess) e = run();
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; e == hipSuccess && r < reps; r++)
        e = run();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / reps;
    *wave_insns = (double)blocks * 4.0 * iters * 8.0 * (kind == P_ADDC ? 2.0 : 1.0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return (int)e;
}
