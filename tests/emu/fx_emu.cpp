// CPU build of the device limb arithmetic (cft-fp256_amd/csrc/fx_*.hpp) with
// the portable fx_hw.hpp beside this file: one element at a time, for the
// not-gpu tests to diff against the oracle.  Test infrastructure only.
#include "fx_arith.hpp"
#include "fx_divsqrt.hpp"
#include "fx_misc.hpp"
#include "hardcase.hpp"

#include <string.h>

using namespace fx;

template <int NL>
static void one(int op, int rnd, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                uint32_t *out, uint32_t *flags)
{
    uint32_t xa[NL], xb[NL], xc[NL], o[NL];
    uint32_t fl = 0;
    for (int i = 0; i < NL; i++) {
        xa[i] = a ? a[i] : 0u;
        xb[i] = b ? b[i] : 0u;
        xc[i] = c ? c[i] : 0u;
    }
    switch (op) {
    case 0: fma_op<NL, OP_FMA>(xa, xb, xc, rnd, o, fl); break;
    case 1: fma_op<NL, OP_ADD>(xa, xb, xc, rnd, o, fl); break;
    case 2: fma_op<NL, OP_SUB>(xa, xb, xc, rnd, o, fl); break;
    case 3: fma_op<NL, OP_MUL>(xa, xb, xc, rnd, o, fl); break;
    case 256: div_op<NL>(xa, xb, rnd, o, fl); break;
    case 257: sqrt_op<NL>(xa, rnd, o, fl); break;
    case 258: add_core<NL>(xa, xc, false, rnd, o, fl); break;   // the two-term ADD path
    case 259: add_core<NL>(xa, xc, true, rnd, o, fl); break;    // and SUB
    default: misc_op<NL>(op, xa, xb, xc, o, fl); break;
    }
    memcpy(out, o, sizeof o);
    *flags = fl;
}

extern "C" int emu_run(int fmt, int op, int rnd, const void *a, const void *b, const void *c,
                       void *d, size_t n, uint32_t *flags_out)
{
    const int nl = 1 << fmt;
    uint32_t acc = 0;
    for (size_t i = 0; i < n; i++) {
        const uint32_t *pa = a ? (const uint32_t *)a + i * nl : nullptr;
        const uint32_t *pb = b ? (const uint32_t *)b + i * nl : nullptr;
        const uint32_t *pc = c ? (const uint32_t *)c + i * nl : nullptr;
        uint32_t *pd = (uint32_t *)d + i * nl;
        uint32_t fl = 0;
        switch (fmt) {
        case 0: one<1>(op, rnd, pa, pb, pc, pd, &fl); break;
        case 1: one<2>(op, rnd, pa, pb, pc, pd, &fl); break;
        case 2: one<4>(op, rnd, pa, pb, pc, pd, &fl); break;
        default: one<8>(op, rnd, pa, pb, pc, pd, &fl); break;
        }
        acc |= fl;
    }
    *flags_out = acc;
    return 0;
}

extern "C" int emu_one(int fmt, int op, int rnd, const void *a, const void *b, const void *c,
                       void *d, uint32_t *flags)
{
    return emu_run(fmt, op, rnd, a, b, c, d, 1, flags);
}

// The two-term add path against the FMA path's ADD / SUB, element by
// element, bits and flags: returns the number of differing elements and
// the first one's index in *first.
extern "C" size_t emu_cmp_add(int fmt, int rnd, int sub, const void *a, const void *c, size_t n,
                              size_t *first)
{
    const int nl = 1 << fmt;
    size_t bad = 0;
    *first = (size_t)-1;
    for (size_t i = 0; i < n; i++) {
        const uint32_t *pa = (const uint32_t *)a + i * nl, *pc = (const uint32_t *)c + i * nl;
        uint32_t d0[8], d1[8], f0 = 0, f1 = 0;
        switch (fmt) {
        case 0: one<1>(sub ? 2 : 1, rnd, pa, nullptr, pc, d0, &f0);
                one<1>(sub ? 259 : 258, rnd, pa, nullptr, pc, d1, &f1); break;
        case 1: one<2>(sub ? 2 : 1, rnd, pa, nullptr, pc, d0, &f0);
                one<2>(sub ? 259 : 258, rnd, pa, nullptr, pc, d1, &f1); break;
        case 2: one<4>(sub ? 2 : 1, rnd, pa, nullptr, pc, d0, &f0);
                one<4>(sub ? 259 : 258, rnd, pa, nullptr, pc, d1, &f1); break;
        default: one<8>(sub ? 2 : 1, rnd, pa, nullptr, pc, d0, &f0);
                 one<8>(sub ? 259 : 258, rnd, pa, nullptr, pc, d1, &f1); break;
        }
        if (memcmp(d0, d1, 4 * (size_t)nl) != 0 || f0 != f1) {
            if (!bad)
                *first = i;
            bad++;
        }
    }
    return bad;
}

// The hard-case generator (hardcase.hpp), host build: cases base .. base+n-1,
// the arrays laid out as cftp_hardcases (lib/libcftprobe.so) writes them on
// the GPU, so the two can be compared bit for bit.
template <int NL>
static void hard(int op, uint64_t seed, uint64_t base, size_t n, uint32_t *a, uint32_t *b,
                 uint32_t *c, uint8_t *fam)
{
    for (size_t i = 0; i < n; i++) {
        uint32_t xa[NL], xb[NL], xc[NL];
        const int f = hc::case_at<NL>(seed, base + i, op, xa, xb, xc);
        memcpy(a + i * NL, xa, sizeof xa);
        if (b)
            memcpy(b + i * NL, xb, sizeof xb);
        if (c)
            memcpy(c + i * NL, xc, sizeof xc);
        if (fam)
            fam[i] = (uint8_t)f;
    }
}

extern "C" int emu_hardcases(int fmt, int op, uint64_t seed, uint64_t base, size_t n, void *a,
                             void *b, void *c, uint8_t *fam)
{
    if (fmt < 0 || fmt > 3 || !hc::op_known(op))
        return -1;
    if (!hc::op_reads_b(op))
        b = nullptr;
    if (!hc::op_reads_c(op))
        c = nullptr;
    uint32_t *A = (uint32_t *)a, *B = (uint32_t *)b, *C = (uint32_t *)c;
    switch (fmt) {
    case 0: hard<1>(op, seed, base, n, A, B, C, fam); break;
    case 1: hard<2>(op, seed, base, n, A, B, C, fam); break;
    case 2: hard<4>(op, seed, base, n, A, B, C, fam); break;
    default: hard<8>(op, seed, base, n, A, B, C, fam); break;
    }
    return 0;
}
