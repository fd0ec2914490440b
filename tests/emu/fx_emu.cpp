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
    // the sequencer's variant of the FMA core (FMA, ADD, SUB, MUL)
    case 260: fma_core<NL, false>(OP_FMA, xa, xb, xc, rnd, o, fl); break;
    case 261: fma_core<NL, false>(OP_ADD, xa, xb, xc, rnd, o, fl); break;
    case 262: fma_core<NL, false>(OP_SUB, xa, xb, xc, rnd, o, fl); break;
    case 263: fma_core<NL, false>(OP_MUL, xa, xb, xc, rnd, o, fl); break;
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

// The square root's digit recurrence (isqrt_limbs, binary256 width) held to
// what fixes it: r^2 <= M < (r+1)^2 and rem_nz == (r^2 != M), exactly, on n
// radicands from a seeded generator - random ones (top bit at 511 or 510)
// and, at three in four, squares s^2 of random 256-bit roots plus -1, 0,
// +1, 2s or 2s-1, some with runs of zero or all-one limbs, where the digit
// estimates sit on integer boundaries.  Returns the number that fail.
static int cmp16(const uint32_t *a, const uint32_t *b)
{
    for (int i = 15; i >= 0; i--)
        if (a[i] != b[i])
            return a[i] < b[i] ? -1 : 1;
    return 0;
}

static uint64_t sm64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" size_t emu_isqrt_check(uint64_t seed, size_t n)
{
    constexpr int NL = 8;
    size_t bad = 0;
    uint64_t st = seed;
    for (size_t k = 0; k < n; k++) {
        uint32_t M[16];
        const int mode = (int)(k % 4);
        if (mode == 0) {
            for (int i = 0; i < 16; i++)
                M[i] = (uint32_t)sm64(st);
            M[15] |= 0x40000000u;
            M[15] = (sm64(st) & 1) ? (M[15] | 0x80000000u) : (M[15] & 0x7FFFFFFFu);
        } else {
            uint32_t s[NL];
            for (int i = 0; i < NL; i++)
                s[i] = (uint32_t)sm64(st);
            s[NL - 1] |= 0x80000000u;
            if (mode == 3) {
                const int z = (int)(sm64(st) % NL);
                for (int i = 0; i < z; i++)
                    s[i] = (sm64(st) & 1) ? 0u : 0xFFFFFFFFu;
            }
            uint32_t sq[16];
            mul_limbs<NL, NL>(s, s, sq);
            uint32_t add[16] = {0};
            int sub1 = 0;
            switch ((int)(sm64(st) % 5)) {
            case 0: sub1 = 1; break;                       // s^2 - 1
            case 1: break;                                 // s^2
            case 2: add[0] = 1; break;                     // s^2 + 1
            default: {                                     // s^2 + 2s (- 1)
                uint64_t c = 0;
                for (int i = 0; i < NL; i++) {
                    const uint64_t v = ((uint64_t)s[i] << 1) + c;
                    add[i] = (uint32_t)v;
                    c = v >> 32;
                }
                add[NL] = (uint32_t)c;
                sub1 = (int)(sm64(st) & 1);
            }
            }
            uint64_t c = 0;
            for (int i = 0; i < 16; i++) {
                c += (uint64_t)sq[i] + add[i];
                M[i] = (uint32_t)c;
                c >>= 32;
            }
            uint64_t br = (uint64_t)sub1;
            for (int i = 0; i < 16; i++) {
                const uint64_t v = (uint64_t)M[i] - br;
                M[i] = (uint32_t)v;
                br = (v >> 63) & 1u;
            }
            if (!(M[15] >> 30))
                M[15] |= 0x40000000u;                      // keep the contract's top bits
        }
        uint32_t r[NL];
        bool nz;
        isqrt_limbs<NL>(M, r, nz);
        uint32_t sq[16], diff[16], two_r[16] = {0};
        mul_limbs<NL, NL>(r, r, sq);
        uint64_t br = 0, c = 0;
        for (int i = 0; i < 16; i++) {
            const uint64_t v = (uint64_t)M[i] - sq[i] - br;
            diff[i] = (uint32_t)v;
            br = (v >> 63) & 1u;
        }
        for (int i = 0; i < NL; i++) {
            const uint64_t v = ((uint64_t)r[i] << 1) + c;
            two_r[i] = (uint32_t)v;
            c = v >> 32;
        }
        two_r[NL] = (uint32_t)c;
        bool zero = true;
        for (int i = 0; i < 16; i++)
            zero = zero && diff[i] == 0;
        if (cmp16(sq, M) > 0 || cmp16(diff, two_r) > 0 || nz == zero)
            bad++;
    }
    return bad;
}
