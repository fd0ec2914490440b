// Counter-based hard cases for the differential tests: the generator, one
// case at a time, as device code (k_hardcase.hip launches it over a chunk)
// that also compiles for the host, so that a CPU build can regenerate any
// case and check that both devices make the same bits.
//
// Case g of seed S is a pure function of (S, g, format, op): a chunk can be
// regenerated from any starting index, and the operands of every case are
// copied back to the host, where the reference's CPU library and the oracle
// compute the expected results from the same bits.  The generator's own
// arithmetic (a rounded product, used to place c against a*b) therefore only
// steers the cases; it is not checked by them.
//
// The families follow the classes the reference's fixed-tier stream aims at
// (host/tools/fx256_stream.h:11-41: raw bits, operands at the subnormal
// edge, products at the underflow and overflow edges, deep and exact
// cancellation, ties, c swept across the product's binades, specials,
// cancellation near EMIN and EMAX, c at the normal boundary, overflow by a
// rounding carry, the alignment edges).  The generator is written for this
// library and is not the reference's: the cases differ, the classes match.
// Family = g mod 14:
//   0  raw bits                      7  c swept over +-(2P+40) binades
//   1  operands at the subnormal edge 8  specials (zeros, infs, NaNs, edges)
//   2  product at the underflow edge  9  cancellation near EMIN
//   3  product at the overflow edge  10  cancellation near EMAX
//   4  deep cancellation c ~ -(a*b)  11  c at the normal boundary
//   5  exact cancellation, and the    12  overflow by a rounding carry
//      product's rounding error       13  alignment edges of c against a*b
//   6  ties and near-ties of a*b + c      (+-31..43, P+-4, 2P, the window T)
// For ADD and SUB the operands are (a*b rounded, c) and (a*b rounded, -c),
// so the sum is the family's a*b + c; MUL takes (a, b); DIV (a*b rounded
// +-2 ulps, b); SQRT (a*a rounded +-2 ulps).  Families 0 and 8 pass their raw
// operands to every op.
#pragma once

#include <stdint.h>

#include "fx_arith.hpp"

namespace hc {

using namespace fx;

constexpr int NFAM = 14;
enum : int { G_FMA = 0, G_ADD = 1, G_SUB = 2, G_MUL = 3, G_DIV = 256, G_SQRT = 257 };

// splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): a 64-bit counter through a mixing function.
struct Rng {
    uint64_t s;
    __device__ uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __device__ uint32_t u32() { return (uint32_t)(next() >> 32); }
    // uniform in [lo, hi] (hi - lo < 2^32)
    __device__ int in(int lo, int hi) { return lo + (int)(u32() % (uint32_t)(hi - lo + 1)); }
    __device__ int bit() { return (int)(next() >> 63); }
};

template <int NL>
__device__ __forceinline__ int bexp(const uint32_t (&x)[NL])
{
    return (int)((x[NL - 1] >> F<NL>::TOPF) & F<NL>::EMASK);
}

template <int NL>
__device__ __forceinline__ void neg(uint32_t (&x)[NL])
{
    x[NL - 1] ^= 0x80000000u;
}

template <int NL>
__device__ __forceinline__ void copy(const uint32_t (&s)[NL], uint32_t (&d)[NL])
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        d[i] = s[i];
}

// Significand kinds: 0 random fraction, 1 all ones, 2 zero (a power of two),
// 3 short (only the top P/2-1 fraction bits random, so a product of two is
// exact), 4 one unit in the last place above a power of two, 5 1.5.
enum : int { K_RAND = 0, K_ONES = 1, K_ZERO = 2, K_SHORT = 3, K_LOW = 4, K_HALF = 5 };

// (-1)^sign * sig * 2^E with the significand of `kind`; E above the format's
// range is clamped to the largest binade, E below EMIN gives the subnormal
// (or the smallest subnormal when everything is shifted out).
template <int NL>
__device__ __forceinline__ void value(Rng &r, int sign, int E, int kind, uint32_t (&x)[NL])
{
    using Fm = F<NL>;
#pragma unroll
    for (int i = 0; i < NL; i++)
        x[i] = (kind == K_RAND || kind == K_SHORT) ? r.u32() : (kind == K_ONES ? ~0u : 0u);
    x[NL - 1] &= Fm::FMASK;
    if (kind == K_SHORT) {
        constexpr int clr = Fm::MANW - (Fm::P / 2 - 1);  // low fraction bits cleared
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const int lo = 32 * i;
            if (lo + 32 <= clr)
                x[i] = 0;
            else if (lo < clr)
                x[i] &= ~((1u << (clr - lo)) - 1u);
        }
    }
    if (kind == K_LOW)
        x[0] |= 1u;
    if (kind == K_HALF)
        x[NL - 1] |= 1u << (Fm::TOPF - 1);
    int be = E + Fm::BIAS;
    if (be > 2 * Fm::BIAS)
        be = 2 * Fm::BIAS;
    if (be >= 1) {
        x[NL - 1] |= (uint32_t)be << Fm::TOPF;
    } else {
        x[NL - 1] |= 1u << Fm::TOPF;              // the implicit bit, then denormalise
        uint32_t jam = 0;
        const int sh = 1 - be;
        shr_jam<NL>(x, sh > 32 * NL ? 32u * NL : (uint32_t)sh, jam);
        if (is_zero<NL>(x))
            x[0] = 1u;
    }
    x[NL - 1] |= (uint32_t)(sign & 1) << 31;
}

template <int NL>
__device__ __forceinline__ void raw(Rng &r, uint32_t (&x)[NL])
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        x[i] = r.u32();
}

// Exponent of the leading bit of a finite nonzero encoding.
template <int NL>
__device__ __forceinline__ int lead_exp(const uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    const int be = bexp<NL>(x);
    if (be)
        return be - Fm::BIAS;
    uint32_t f[NL];
    copy<NL>(x, f);
    f[NL - 1] &= Fm::FMASK;
    return Fm::EMIN - Fm::MANW + bitlen<NL>(f) - 1;
}

// The magnitude's encoding plus k (signed): k units in the last place,
// crossing binades as the encoding does; clamped at zero and at infinity.
template <int NL>
__device__ __forceinline__ void add_ulps(uint32_t (&x)[NL], int k)
{
    using Fm = F<NL>;
    const uint32_t sgn = x[NL - 1] & 0x80000000u;
    x[NL - 1] &= 0x7fffffffu;
    uint64_t c = (uint32_t)(k >= 0 ? k : -k);
    if (k >= 0) {
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const uint64_t t = (uint64_t)x[i] + c;
            x[i] = (uint32_t)t;
            c = t >> 32;
        }
    } else {
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const uint64_t t = (uint64_t)x[i] - c;
            x[i] = (uint32_t)t;
            c = t >> 63;
        }
        if (c) {
#pragma unroll
            for (int i = 0; i < NL; i++)
                x[i] = 0;
        }
    }
    if ((x[NL - 1] >> Fm::TOPF) >= Fm::EMASK) {   // reached (or passed) infinity
#pragma unroll
        for (int i = 0; i < NL; i++)
            x[i] = 0;
        x[NL - 1] = Fm::EMASK << Fm::TOPF;
    }
    x[NL - 1] |= sgn;
}

template <int NL>
__device__ __forceinline__ void product(const uint32_t (&a)[NL], const uint32_t (&b)[NL],
                                        uint32_t (&p)[NL])
{
    uint32_t z[NL], fl = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        z[i] = 0;
    fma_core<NL>(OP_MUL, a, b, z, RNE, p, fl);
}

// Every draw below is a statement of its own (or the one draw of an
// expression): the order of evaluation of a call's arguments is unspecified
// in C++, and host and device compilers do choose differently, so two draws
// in one argument list would make the CPU and GPU builds different cases.

// A random sign, then value().
template <int NL>
__device__ __forceinline__ void rvalue(Rng &r, int E, int kind, uint32_t (&x)[NL])
{
    const int s = r.bit();
    value<NL>(r, s, E, kind, x);
}

__device__ __forceinline__ int any_kind(Rng &r)
{
    const int k = r.in(0, 7);
    return k <= 2 ? K_RAND : k - 2;               // random three times in eight
}

// a and b whose product leads at about Et (ea + eb = Et; the product of two
// significands adds 0 or 1), spread so neither is extreme.
template <int NL>
__device__ __forceinline__ void pair_at(Rng &r, int Et, int ka, int kb, uint32_t (&a)[NL],
                                        uint32_t (&b)[NL])
{
    using Fm = F<NL>;
    constexpr int spread = Fm::EMAX / 2 < 1000 ? Fm::EMAX / 2 : 1000;
    const int ea = Et / 2 + r.in(-spread, spread);
    rvalue<NL>(r, ea, ka, a);
    rvalue<NL>(r, Et - ea, kb, b);
}

// pair_at with both kinds drawn by any_kind, a's first.
template <int NL>
__device__ __forceinline__ void pair_any(Rng &r, int Et, uint32_t (&a)[NL], uint32_t (&b)[NL])
{
    const int ka = any_kind(r);
    const int kb = any_kind(r);
    pair_at<NL>(r, Et, ka, kb, a, b);
}

template <int NL>
__device__ __forceinline__ void signed_zero(int s, uint32_t (&x)[NL])
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        x[i] = 0;
    x[NL - 1] = (uint32_t)(s & 1) << 31;
}

// An operand at the subnormal edge: subnormal, just above EMIN, a few bits
// above the smallest subnormal, or near 1 (so a product with a subnormal
// stays at the edge).
template <int NL>
__device__ __forceinline__ void edge_operand(Rng &r, uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    const int w = r.in(0, 3);
    int E;
    switch (w) {
    case 0: E = Fm::EMIN - r.in(1, Fm::P); break;
    case 1: E = Fm::EMIN + r.in(0, 2); break;
    case 2: E = Fm::EMIN - Fm::P + r.in(1, 8); break;
    default: E = r.in(-Fm::P, Fm::P); break;
    }
    const int k = w == 2 ? (int)K_RAND : any_kind(r);
    rvalue<NL>(r, E, k, x);
}

template <int NL>
__device__ __forceinline__ void special(Rng &r, uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    const int s = r.bit();
    const int w = r.in(0, 13);
    switch (w) {
    case 0: signed_zero<NL>(s, x); break;
    case 1: case 2:
#pragma unroll
        for (int i = 0; i < NL; i++) x[i] = 0;
        x[NL - 1] = (Fm::EMASK << Fm::TOPF) | ((uint32_t)s << 31);   // +-inf
        break;
    case 3: raw<NL>(r, x);                        // quiet NaN, random payload
        x[NL - 1] |= (Fm::EMASK << Fm::TOPF) | (1u << (Fm::TOPF - 1));
        break;
    case 4: raw<NL>(r, x);                        // signalling NaN, nonzero payload
        x[NL - 1] = (x[NL - 1] & (0x80000000u | (Fm::FMASK >> 1))) | (Fm::EMASK << Fm::TOPF);
        x[0] |= 1u;
        break;
    case 5: value<NL>(r, s, Fm::EMIN - Fm::MANW, K_ZERO, x); break;   // smallest subnormal
    case 6: value<NL>(r, s, Fm::EMIN - 1, K_ONES, x); break;          // largest subnormal
    case 7: value<NL>(r, s, Fm::EMIN, K_ZERO, x); break;              // smallest normal
    case 8: value<NL>(r, s, Fm::EMAX, K_ONES, x); break;              // largest finite
    case 9: value<NL>(r, s, 0, K_ZERO, x); break;                     // +-1
    case 10: value<NL>(r, s, 0, K_LOW, x); break;                     // 1 + ulp
    case 11: value<NL>(r, s, -1, K_ONES, x); break;                   // 1 - ulp/2
    case 12: {
        const int E = r.in(-64, 64);
        value<NL>(r, s, E, K_RAND, x);
        break;
    }
    default: raw<NL>(r, x); break;
    }
}

template <int NL>
__device__ __forceinline__ void make_case(Rng &r, int fam, int op, uint32_t (&a)[NL],
                                          uint32_t (&b)[NL], uint32_t (&c)[NL])
{
    using Fm = F<NL>;
    constexpr int P = Fm::P, EMIN = Fm::EMIN, EMAX = Fm::EMAX;
    uint32_t p[NL];
    bool have_p = false;
    switch (fam) {
    case 0:
        raw<NL>(r, a);
        raw<NL>(r, b);
        raw<NL>(r, c);
        break;
    case 1: {                                     // operands at the subnormal edge
        edge_operand<NL>(r, a);
        edge_operand<NL>(r, b);
        const int w = r.in(0, 3);
        if (w == 3) {
            const int s = r.bit();
            signed_zero<NL>(s, c);
        } else {
            const int E = w == 0 ? EMIN - r.in(0, P) : EMIN + r.in(0, 2 * P);
            const int k = any_kind(r);
            rvalue<NL>(r, E, k, c);
        }
        break;
    }
    case 2: {                                     // product at the underflow edge
        const int Et = EMIN + r.in(-(P + 3), 3);
        pair_any<NL>(r, Et, a, b);
        const int w = r.in(0, 3);
        if (w == 0) {
            const int s = r.bit();
            signed_zero<NL>(s, c);
        } else if (w == 3) {
            rvalue<NL>(r, EMIN, K_ZERO, c);
            const int u = r.in(-3, 3);
            add_ulps<NL>(c, u);
        } else {
            const int E = w == 1 ? EMIN - r.in(1, P) : Et + r.in(-3, 3);
            const int k = any_kind(r);
            rvalue<NL>(r, E, k, c);
        }
        break;
    }
    case 3: {                                     // product at the overflow edge
        const int Et = EMAX + r.in(-2, 1);
        const int ka = r.bit() ? K_ONES : K_RAND;
        const int kb = any_kind(r);
        pair_at<NL>(r, Et, ka, kb, a, b);
        const int w = r.in(0, 3);
        if (w == 0) {
            const int s = r.bit();
            signed_zero<NL>(s, c);
        } else if (w == 2) {
            rvalue<NL>(r, EMAX, K_ONES, c);
        } else {
            const int E = w == 1 ? EMAX - r.in(0, P + 2) : r.in(-10, 10);
            const int k = w == 1 ? any_kind(r) : (int)K_RAND;
            rvalue<NL>(r, E, k, c);
        }
        break;
    }
    case 4: case 9: case 10: {                    // cancellation: c = -(a*b rounded) +- k ulps
        const int Et = fam == 4 ? r.in(EMIN / 2, EMAX / 2)
                     : fam == 9 ? EMIN + r.in(-4, P + 4) : EMAX - r.in(1, 4);
        pair_any<NL>(r, Et, a, b);
        product<NL>(a, b, p);
        have_p = true;
        copy<NL>(p, c);
        neg<NL>(c);
        const int u = r.in(-8, 8);
        add_ulps<NL>(c, u);
        break;
    }
    case 5: {                                     // exact cancellation / the product's error
        const int Et = r.in(EMIN + 2 * P, EMAX - 2);
        const int k = r.bit() ? K_SHORT : K_RAND;
        pair_at<NL>(r, Et, k, k, a, b);
        product<NL>(a, b, p);
        have_p = true;
        copy<NL>(p, c);
        neg<NL>(c);
        break;
    }
    case 6: {                                     // ties of an exact product and c
        const int Et = r.in(EMIN / 2, EMAX / 2);
        pair_at<NL>(r, Et, K_SHORT, K_SHORT, a, b);
        product<NL>(a, b, p);
        have_p = true;
        const int Ep = lead_exp<NL>(p);
        const int w = r.in(0, 3);
        switch (w) {
        case 0: rvalue<NL>(r, Ep - P, K_ZERO, c); break;        // half an ulp
        case 1: rvalue<NL>(r, Ep - P, K_LOW, c); break;         // just above
        case 2: rvalue<NL>(r, Ep - P - 1, K_ONES, c); break;    // just below
        default: rvalue<NL>(r, Ep - P, K_HALF, c); break;       // three quarters
        }
        break;
    }
    case 7: {                                     // c swept across the product's binades
        const int Et = r.in(EMIN / 4, EMAX / 4);
        pair_any<NL>(r, Et, a, b);
        const int E = Et + r.in(-(2 * P + 40), 2 * P + 40);
        const int k = any_kind(r);
        rvalue<NL>(r, E, k, c);
        break;
    }
    case 8:
        special<NL>(r, a);
        special<NL>(r, b);
        special<NL>(r, c);
        break;
    case 11: {                                    // c at the normal boundary
        const int E = r.bit() ? EMIN : EMIN - 1;
        const int k = r.bit() ? K_ZERO : K_ONES;
        rvalue<NL>(r, E, k, c);
        const int u = r.in(-3, 3);
        add_ulps<NL>(c, u);
        const int Et = r.bit() ? EMIN + r.in(-3, 3) : r.in(EMIN / 2, EMAX / 2);
        pair_any<NL>(r, Et, a, b);
        break;
    }
    case 12: {                                    // overflow by a rounding carry
        int Ep;
        if (r.bit()) {                            // all-ones significands: a*b ~ 4 * 2^(ea+eb)
            Ep = EMAX - r.in(0, 1);
            pair_at<NL>(r, Ep - 1, K_ONES, K_ONES, a, b);
        } else {                                  // the largest finite times ~1
            Ep = EMAX;
            rvalue<NL>(r, EMAX, K_ONES, a);
            const int w = r.in(0, 2);
            rvalue<NL>(r, w == 2 ? -1 : 0, w == 0 ? K_ZERO : w == 1 ? K_LOW : K_ONES, b);
        }
        const int w = r.in(0, 2);
        // c pushes the product up (its sign) three times in four
        const int s = r.in(0, 3) ? (int)((a[NL - 1] ^ b[NL - 1]) >> 31) : r.bit();
        value<NL>(r, s, w == 0 ? Ep - P - 1 : Ep - P, w == 0 ? K_ONES : w == 1 ? K_ZERO : K_LOW,
                  c);
        break;
    }
    default: {                                    // 13: alignment edges
        const int Et = r.in(EMIN / 4, EMAX / 4);
        pair_any<NL>(r, Et, a, b);
        product<NL>(a, b, p);
        have_p = true;
        const int Ep = lead_exp<NL>(p);
        const int s = r.bit() ? 1 : -1;
        const int w = r.in(0, 7);
        int d;
        switch (w) {
        case 0: d = r.in(31, 43); break;
        case 1: d = -(P + r.in(31, 43)); break;
        case 2: d = s * (2 * P + r.in(-4, 8)); break;
        case 3: d = s * (P + r.in(-4, 4)); break;
        case 4: d = r.in(-3, 3); break;
        case 5: d = s * (Fm::T + r.in(-4, 4)); break;
        case 6: d = s * (Fm::T - P + r.in(-4, 4)); break;
        default: d = s * (Fm::T - 2 * P + r.in(-4, 4)); break;
        }
        const int k = r.in(0, 3);
        rvalue<NL>(r, Ep + d, k == 0 ? K_RAND : k == 1 ? K_ONES : k == 2 ? K_ZERO : K_LOW, c);
        break;
    }
    }

    // the op's operands from the family's (a, b, c)
    const bool rawfam = fam == 0 || fam == 8;
    if (op == G_ADD || op == G_SUB || op == G_DIV) {
        if (!rawfam) {
            if (!have_p)
                product<NL>(a, b, p);
            if (op == G_DIV) {
                const int u = r.in(-2, 2);
                add_ulps<NL>(p, u);
                copy<NL>(p, a);
            } else {
                copy<NL>(p, a);
                if (op == G_SUB)
                    neg<NL>(c);
            }
        }
    } else if (op == G_SQRT && !rawfam) {
        uint32_t q[NL];
        product<NL>(a, a, q);
        const int u = r.in(-2, 2);
        add_ulps<NL>(q, u);
        copy<NL>(q, a);
    }
}

// Case g of `seed` for op: its operands (those op does not read are left as
// the family made them) and its family.
template <int NL>
__device__ __forceinline__ int case_at(uint64_t seed, uint64_t g, int op, uint32_t (&a)[NL],
                                       uint32_t (&b)[NL], uint32_t (&c)[NL])
{
    Rng r{seed ^ (0xD1B54A32D192ED03ull * (g + 1)) ^ ((uint64_t)(op & 0x1ff) << 52) ^
          ((uint64_t)NL << 61)};
    r.next();
    const int fam = (int)(g % NFAM);
    make_case<NL>(r, fam, op, a, b, c);
    return fam;
}

inline bool op_known(int op)
{
    return op == G_FMA || op == G_ADD || op == G_SUB || op == G_MUL || op == G_DIV ||
           op == G_SQRT;
}
inline bool op_reads_b(int op) { return op == G_FMA || op == G_MUL || op == G_DIV; }
inline bool op_reads_c(int op) { return op == G_FMA || op == G_ADD || op == G_SUB; }

} // namespace hc
