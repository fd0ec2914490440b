// cftx_reduce_noalloc: cft_reduce over operands already in device memory,
// for callers that cannot allocate (code that a HIP graph may capture around
// it, tight serving loops, harnesses that forbid hipMalloc).
//
// cft_reduce keeps the tree's partial results in growable device arenas of
// the handle.  This variant keeps them in `scratch`, one array of device
// memory per GPU that the library holds until the process ends, so a call
// allocates nothing: no hipMalloc, no hipHostMalloc, no host heap.  Its
// layout:
//   [0, 8 KiB)           the flag slots (64 lines of 128 bytes, kernels.h),
//                        zero between calls: the gather launch that reads
//                        them zeroes them
//   [8 KiB, +256)        sumSquare / sumAbs: the flag word and the word
//                        recording an infinity or a signalling NaN, cleared
//                        by a fill launch at the start of every call
//   [8 KiB + 256, ...)   the partials: ceil(n/16) nodes of the first launch,
//                        then ceil(n/256) of the second, the later launches
//                        reusing the first area
// The partial area is 8.5 MiB, so n is at most 2^22 binary256, 2^23
// binary128, 2^24 binary64 or 2^25 binary32 elements; larger n is refused.
//
// Every call on a GPU shares that GPU's array: two calls that overlap on one
// GPU (two handles, two threads) race on it and their results are undefined.
// Calls on one handle, which is not thread-safe anyway, are ordered on its
// stream.
#include "cft_internal.hpp"
#include "kernels.h"

#include <mutex>

namespace {

constexpr size_t kSlotBytes = (size_t)CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS * 4;
constexpr size_t kWordOff = kSlotBytes;
constexpr size_t kPartOff = kSlotBytes + 256;
constexpr size_t kPartBytes = (size_t)17 << 19;      // 8.5 MiB
constexpr size_t kScratchBytes = kPartOff + kPartBytes;
constexpr int kMaxDevices = 64;

const int kExpBits[4] = {8, 11, 15, 19};

// The result's last step, one thread: mode 0 copies the root node to d;
// mode 1 does the same and applies 754-2019 9.4 for sumSquare / sumAbs (a
// NaN over an array holding an infinity is +inf, INVALID exactly when the
// array holds a signalling NaN; words[1] bit 0 records an infinity, bit 1 a
// signalling NaN, words[0] holds the flags); modes 2 and 3 store the empty
// array's identity, +0 or -inf.
__global__ void k_noalloc_finish(const uint32_t *src, uint32_t *d, uint32_t *words, int nl,
                                 int ew, int mode)
{
    uint32_t x[8];
    const uint32_t emask = ((1u << ew) - 1u) << (31 - ew);
    for (int k = 0; k < nl; k++)
        x[k] = mode >= 2 ? 0u : src[k];
    if (mode == 3)
        x[nl - 1] = 0x80000000u | emask;
    if (mode == 1 && (words[1] & 1u)) {
        bool frac = (x[nl - 1] & ~(0x80000000u | emask)) != 0;
        for (int k = 0; k < nl - 1; k++)
            frac |= x[k] != 0;
        if ((x[nl - 1] & emask) == emask && frac) {
            for (int k = 0; k < nl - 1; k++)
                x[k] = 0;
            x[nl - 1] = emask;
            words[0] = (words[1] & 2u) ? (uint32_t)CFT_FLAG_INVALID : 0u;
        }
    }
    for (int k = 0; k < nl; k++)
        d[k] = x[k];
}

} // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_mu;
uint8_t *g_scratch[kMaxDevices];     // each GPU's address of `scratch`

// The device address of `scratch` on dev's GPU (current), looked up once
// per GPU; its flag slots and words are cleared then.
cft_status scratch_of(cft_device *dev, uint8_t **out, const char *who)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const int g = dev->hip_index;
    if (g < 0 || g >= kMaxDevices) {
        cft_set_error("%s: GPU %d is past the %d this entry point keeps scratch for", who, g,
                      kMaxDevices);
        return CFT_ERR_UNSUPPORTED;
    }
    if (!g_scratch[g]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            cft_set_error("%s: the scratch array's address on GPU %d could not be found (%s)",
                          who, g, hipGetErrorString(e));
            return CFT_ERR_INTERNAL;
        }
        e = cftk_fill32((uint32_t *)p, 0u, kPartOff / 4, dev->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "clearing the scratch array's flag words");
        g_scratch[g] = (uint8_t *)p;
    }
    *out = g_scratch[g];
    return CFT_OK;
}

} // namespace

extern "C" CFT_API cft_status cftx_reduce_noalloc(cft_device *dev, cft_op op, cft_format fmt,
                                                  cft_round rnd, const void *a, const void *b,
                                                  void *d, size_t n, uint32_t *flags_out,
                                                  uint32_t *bus_out)
{
    static const char *who = "cftx_reduce_noalloc";
    cft_clear_error();
    if (bus_out)
        *bus_out = 0;
    cft_status st = reduce_check(who, dev, (int)op, (int)fmt, (int)rnd, a, b, d, n);
    if (st != CFT_OK)
        return st;
    const size_t esz = cft_format_size(fmt);
    const size_t first = (n + 15) / 16, second = (first + 15) / 16;
    if ((first + second) * esz > kPartBytes) {
        cft_set_error("%s: n = %zu elements of %zu bytes need %zu bytes of partial results "
                      "and the scratch array holds %zu (at most %zu elements of this format)",
                      who, n, esz, (first + second) * esz, kPartBytes,
                      ((size_t)1 << 27) / esz);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const bool o94 = op == CFT_SUMSQ || op == CFT_SUMABS;
    const size_t align = esz < 16 ? esz : 16;
    if ((st = dev_enter(dev, who)) != CFT_OK ||
        (st = device_window(who, dev, "d", d, esz, align)) != CFT_OK ||
        (n && (st = device_window(who, dev, "a", a, n * esz, align)) != CFT_OK) ||
        (n && op == CFT_DOT && (st = device_window(who, dev, "b", b, n * esz, align)) != CFT_OK))
        return st;
    uint8_t *S = nullptr;
    if ((st = scratch_of(dev, &S, who)) != CFT_OK)
        return st;
    uint32_t *slots = (uint32_t *)S, *words = (uint32_t *)(S + kWordOff);
    void *part[2] = {S + kPartOff, S + kPartOff + first * esz};
    const int nl = (int)(esz / 4), ew = kExpBits[fmt];
    hipStream_t sm = dev->stream;
    hipError_t e = hipSuccess;
    if (n == 0) {
        hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(1), 0, sm, (const uint32_t *)nullptr,
                           (uint32_t *)d, words, nl, ew, op == CFT_MAXALL ? 3 : 2);
        e = hipGetLastError();
    } else {
        if (o94)
            e = cftk_fill32(words, 0u, 2, sm);
        const int leaf = op == CFT_DOT ? 1 : op == CFT_SUMABS ? 2 : op == CFT_SUMSQ ? 3 : 0;
        const int kind = op == CFT_MAXALL ? 1 : 0;
        size_t m = n;
        const void *src = a;
        int out = 0;
        for (int pass = 0; e == hipSuccess; pass++) {
            const size_t m2 = (m + 15) / 16;
            e = cftk_reduce((int)fmt, src, pass == 0 ? b : nullptr, part[out], 1, m, m2,
                            pass == 0 ? leaf : 0, kind, (int)rnd, o94 ? 1 : 0,
                            o94 ? words : slots, (pass == 0 && o94) ? words + 1 : nullptr,
                            grid_for(dev, m2), sm);
            src = part[out];
            if (m2 == 1)
                break;
            out ^= 1;
            m = m2;
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(1), 0, sm, (const uint32_t *)src,
                               (uint32_t *)d, words, nl, ew, o94 ? 1 : 0);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = o94 ? cftk_flags_gather(words, 1, dev->hwords_dev, 1, sm)
                    : cftk_flags_gather(slots, CFTK_FLAG_SLOTS, dev->hwords_dev, 1, sm);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(sm);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "running the reduction");
    cft_emit_flags(dev, n ? ((volatile uint32_t *)dev->hwords)[0] : 0u, flags_out);
    return CFT_OK;
}
