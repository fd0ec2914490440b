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
