// The eight-lane binary256 FMA (fx_lp.hpp) as an elementwise kernel and as a
// dependent chain, for its tests and its latency.  Built into
// lib/libcftprobe.so, a test-support library beside libcft.so (which does not
// contain it).
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang diagnostic ignored "-Wpass-failed"

#include "fx_lp.hpp"

using namespace fx;

namespace {

// op 0 FMA, 1 ADD, 2 SUB, 3 MUL, steered to the FMA each is defined as
__device__ __forceinline__ void steer(int op, uint32_t l, uint32_t a7, uint32_t b7, uint32_t &b,
                                      uint32_t &c)
{
    if (op == OP_ADD || op == OP_SUB) {
        b = l == 7 ? (uint32_t)F<8>::BIAS << F<8>::TOPF : 0u;
        if (op == OP_SUB && l == 7)
            c ^= 0x80000000u;
    } else if (op == OP_MUL) {
        c = l == 7 ? (a7 ^ b7) & 0x80000000u : 0u;
    }
}

// element e on lanes 8e .. 8e+7 of the grid (n elements, 8n lanes; a whole
// group of 8 is always inside the grid since the grid is a multiple of 64)
__global__ __launch_bounds__(256) void k_lp_run(int op, int rnd, const uint32_t *__restrict__ A,
                                                const uint32_t *__restrict__ B,
                                                const uint32_t *__restrict__ C,
                                                uint32_t *__restrict__ D, size_t n,
                                                uint8_t *__restrict__ flags_out,
                                                uint8_t *__restrict__ fast_out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t e = t >> 3;
    const uint32_t l = (uint32_t)(t & 7);
    const bool live = e < n;                       // uniform over a group
    const size_t ee = live ? e : 0;
    uint32_t xa = A[ee * 8 + l];
    uint32_t xb = op == OP_ADD || op == OP_SUB ? 0u : B[ee * 8 + l];
    uint32_t xc = op == OP_MUL ? 0u : C[ee * 8 + l];
    const uint32_t a7 = A[ee * 8 + 7], b7 = op == OP_ADD || op == OP_SUB ? 0u : B[ee * 8 + 7];
    steer(op, l, a7, b7, xb, xc);
    uint32_t out = 0, fl = 0;
    const bool ok = lp::fma8(xa, xb, xc, rnd, out, fl);
    if (!ok) {
        // the whole operation on the whole operands, as the one-lane kernels
        uint32_t ra[8], rb[8], rc[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            ra[k] = A[ee * 8 + k];
            rb[k] = op == OP_ADD || op == OP_SUB ? 0u : B[ee * 8 + k];
            rc[k] = op == OP_MUL ? 0u : C[ee * 8 + k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            steer(op, (uint32_t)k, ra[7], rb[7], rb[k], rc[k]);
        fl = 0;
        fma_core<8>(OP_FMA, ra, rb, rc, rnd, o, fl);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if ((uint32_t)k == l)
                out = o[k];
    }
    if (live) {
        D[e * 8 + l] = out;
        if (l == 0) {
            flags_out[e] = (uint8_t)fl;
            if (fast_out)
                fast_out[e] = ok ? 1 : 0;
        }
    }
}

// one wavefront, every group the same chain x <- x*a + c
__global__ __launch_bounds__(64) void k_lp_chain(uint32_t *io, int iters)
{
    const uint32_t l = threadIdx.x & 7;
    uint32_t x = io[l], a = io[8 + l], c = io[16 + l];
    uint32_t fl = 0, slow = 0;
    for (int it = 0; it < iters; it++) {
        uint32_t o, f;
        if (!lp::fma8(x, a, c, RNE, o, f)) {
            slow++;
            o = x;
            f = 0;
        }
        x = o;
        fl |= f;
    }
    if (threadIdx.x < 8) {
        io[24 + l] = x;
        if (l == 0) {
            io[32] = fl;
            io[33] = slow;
        }
    }
}

} // namespace

// n elements of op (0 FMA, 1 ADD, 2 SUB, 3 MUL) in binary256, rounding rnd,
// from device arrays a, b, c (the op's, others may be NULL) into d; each
// element's flags into flags (n bytes) and, where fast is not NULL, 1 for an
// element the eight-lane fast path finished, 0 for one it handed back.
// Synchronous; returns the HIP error code, or -1 for a refused argument.
extern "C" __attribute__((visibility("default"))) int
cftp_lp_run(int op, int rnd, const void *a, const void *b, const void *c, void *d, size_t n,
            uint8_t *flags, uint8_t *fast)
{
    if (op < 0 || op > 3 || rnd < 0 || rnd > 4 || !a || !d || !flags)
        return -1;
    if ((op == 0 || op == 3) && !b)
        return -1;
    if (op != 3 && !c)
        return -1;
    if (n == 0)
        return 0;
    if (n > ((size_t)1 << 28))
        return -1;
    const size_t lanes = n * 8;
    const size_t blocks = (lanes + 255) / 256;
    hipLaunchKernelGGL(k_lp_run, dim3((uint32_t)blocks), dim3(256), 0, 0, op, rnd,
                       (const uint32_t *)a, (const uint32_t *)b, (const uint32_t *)c,
                       (uint32_t *)d, n, flags, fast);
    hipError_t e = hipGetLastError();
    return (int)(e != hipSuccess ? e : hipDeviceSynchronize());
}

// A dependent chain of `iters` eight-lane FMAs x <- x*a + c (RNE) from the 24
// words in (x, a, c); out gets x (8 words), the flags and the count of
// operations the fast path handed back (kept as x); *us_per_op the kernel
// time (HIP events, after one untimed launch of 16) over iters.
extern "C" __attribute__((visibility("default"))) int
cftp_lp_chain_latency(int iters, const uint32_t *in, uint32_t *out, double *us_per_op)
{
    if (iters < 1 || !in || !out || !us_per_op)
        return (int)hipErrorInvalidValue;
    uint32_t *io = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&io, 34 * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpy(io, in, 24 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_lp_chain, dim3(1), dim3(64), 0, 0, io, 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(io, in, 24 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_lp_chain, dim3(1), dim3(64), 0, 0, io, iters);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(out, io + 24, 10 * 4, hipMemcpyDeviceToHost);
    *us_per_op = (double)t * 1e3 / iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (io) (void)hipFree(io);
    return (int)e;
}
