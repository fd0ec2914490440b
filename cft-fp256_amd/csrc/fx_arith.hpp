// Fixed-limb IEEE 754 arithmetic for CDNA4 (gfx950), one lane per element.
//
// Every value lives in registers as NL little-endian 32-bit limbs: NL = 1, 2,
// 4, 8 for binary32/64/128/256.  All loops are over compile-time limb counts
// and fully unrolled, so no array ever leaves the VGPR file; variable shifts
// are logarithmic select networks (v_cndmask) plus one v_alignbit pass, never
// indexed register access.
//
// The operations are the reference contract's, bit for bit and flag for flag:
//   unpack / specials      host/src/softfloat.c:85-116, 312-368
//   rounding               host/src/softfloat.c:120-261 (tininess after
//                          rounding; UNDERFLOW iff tiny and inexact, or the
//                          result rounds to zero; OVERFLOW always inexact)
//   fma and its steering   host/src/softfloat.c:324-468
// The method is this file's own (the reference bounds its alignment with a
// near/far split over a 1200-bit bignum; see fma_op below).
#pragma once

#include <fx_hw.hpp>

namespace fx {

// Exception flags and rounding attributes: host/include/cft.h:481-501.
enum : uint32_t { FL_INV = 1, FL_DZ = 2, FL_OF = 4, FL_UF = 8, FL_NX = 16 };
enum : int { RNE = 0, RTZ = 1, RDN = 2, RUP = 3, RMM = 4 };
// Opcodes: host/include/cft.h:325-472.
enum : int { OP_FMA = 0, OP_ADD = 1, OP_SUB = 2, OP_MUL = 3 };

template <int NL> struct ExpW;
template <> struct ExpW<1> { static constexpr int v = 8; };
template <> struct ExpW<2> { static constexpr int v = 11; };
template <> struct ExpW<4> { static constexpr int v = 15; };
template <> struct ExpW<8> { static constexpr int v = 19; };

// Format constants (host/src/softfloat.c:10-16).
template <int NL> struct F {
    static constexpr int EXPW = ExpW<NL>::v;
    static constexpr int W = 32 * NL;
    static constexpr int MANW = W - 1 - EXPW;
    static constexpr int P = MANW + 1;               // precision
    static constexpr int BIAS = (1 << (EXPW - 1)) - 1;
    static constexpr int EMAX = BIAS;
    static constexpr int EMIN = 1 - BIAS;
    static constexpr int TOPF = MANW - 32 * (NL - 1); // fraction bits, top limb
    static constexpr uint32_t EMASK = (1u << EXPW) - 1u;
    static constexpr uint32_t FMASK = (1u << TOPF) - 1u;
    // FMA window: jam bit, 2P product bits, P + 2 bits of addend headroom
    // above the product, a carry bit.
    static constexpr int WL = (3 * P + 4 + 31) / 32;
};

// ---- limb primitives ------------------------------------------------------

__device__ __forceinline__ int clz32(uint32_t x)
{
    return x ? __builtin_clz(x) : 32;
}

// p[0 .. NA+NB) = x[0 .. NA) * y[0 .. NB), product scanning (column sums).
template <int NA, int NB>
__device__ __forceinline__ void mul_limbs(const uint32_t (&x)[NA], const uint32_t (&y)[NB],
                                          uint32_t (&p)[NA + NB])
{
    uint64_t acc = 0;
    uint32_t hi = 0;
#pragma unroll
    for (int k = 0; k < NA + NB - 1; k++) {
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int j = k - i;
            if (j < 0 || j >= NB)
                continue;
            mac(acc, hi, x[i], y[j]);
        }
        p[k] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)hi << 32);
        hi = 0;
    }
    p[NA + NB - 1] = (uint32_t)acc;
}

// Bit length of an N-limb value (0 for zero).
template <int N>
__device__ __forceinline__ int bitlen(const uint32_t (&x)[N])
{
    int L = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        L = x[i] ? (32 * i + 32 - __builtin_clz(x[i])) : L;
    return L;
}

template <int N>
__device__ __forceinline__ bool is_zero(const uint32_t (&x)[N])
{
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        o |= x[i];
    return o == 0;
}

// x <<= 32*k limbs, zero fill; 0 <= k < 32.
template <int N>
__device__ __forceinline__ void limb_shl(uint32_t (&x)[N], uint32_t k)
{
#pragma unroll
    for (int t = 16; t >= 1; t >>= 1) {
        const bool s = (k & (uint32_t)t) != 0;
#pragma unroll
        for (int i = N - 1; i >= 0; i--)
            x[i] = s ? (i >= t ? x[i - t] : 0u) : x[i];
    }
}

// x >>= 32*k limbs, the dropped limbs ORed into jam; 0 <= k < 32.
template <int N>
__device__ __forceinline__ void limb_shr_jam(uint32_t (&x)[N], uint32_t k, uint32_t &jam)
{
#pragma unroll
    for (int t = 16; t >= 1; t >>= 1) {
        const bool s = (k & (uint32_t)t) != 0;
        uint32_t dropped = 0;
#pragma unroll
        for (int i = 0; i < (t < N ? t : N); i++)
            dropped |= x[i];
        jam |= s ? dropped : 0u;
#pragma unroll
        for (int i = 0; i < N; i++)
            x[i] = s ? (i + t < N ? x[i + t] : 0u) : x[i];
    }
}

// x >>= r bits (any r >= 0), every bit shifted out ORed into jam (nonzero
// iff something was dropped).
template <int N>
__device__ __forceinline__ void shr_jam(uint32_t (&x)[N], uint32_t r, uint32_t &jam)
{
    if (r > 32u * N)
        r = 32u * N;
    const uint32_t k = r >> 5, b = r & 31u;
    limb_shr_jam<N>(x, k, jam);
    jam |= x[0] & ((1u << b) - 1u);
#pragma unroll
    for (int i = 0; i < N; i++)
        x[i] = alignbit(i + 1 < N ? x[i + 1] : 0u, x[i], b);
}

// x <<= s bits, 0 <= s < 32*N, zero fill.
template <int N>
__device__ __forceinline__ void shl(uint32_t (&x)[N], uint32_t s)
{
    const uint32_t k = (s + 31u) >> 5;          // s = 32k - b, 0 <= b < 32
    const uint32_t b = (32u * k - s) & 31u;
    uint32_t y[N + 1];
#pragma unroll
    for (int i = 0; i < N; i++)
        y[i] = x[i];
    y[N] = 0;
    limb_shl<N + 1>(y, k);
#pragma unroll
    for (int i = 0; i < N; i++)
        x[i] = alignbit(y[i + 1], y[i], b);
}

// ---- encodings ------------------------------------------------------------

template <int NL> struct Unp {
    uint32_t m[NL];     // significand, normalised: bit P-1 set when nonzero
    int e;              // exponent of m's lsb
    int sign;
    bool zero, inf, nan, snan;
};

// host/src/softfloat.c:85-116, then a subnormal significand normalised so
// its leading bit is at P-1 (its exponent lowered to match).
template <int NL>
__device__ __forceinline__ void unpack(const uint32_t (&x)[NL], Unp<NL> &u)
{
    using Fm = F<NL>;
    const uint32_t top = x[NL - 1];
    const uint32_t ef = (top >> Fm::TOPF) & Fm::EMASK;
    uint32_t fnz = top & Fm::FMASK;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        fnz |= x[i];
    u.sign = (int)(top >> 31);
    u.nan = ef == Fm::EMASK && fnz != 0;
    u.snan = u.nan && !((top >> (Fm::TOPF - 1)) & 1u);
    u.inf = ef == Fm::EMASK && fnz == 0;
    u.zero = ef == 0 && fnz == 0;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        u.m[i] = x[i];
    u.m[NL - 1] = (top & Fm::FMASK) | (ef ? (1u << Fm::TOPF) : 0u);
    u.e = (ef ? (int)ef : 1) - Fm::BIAS - Fm::MANW;
    if (ef == 0 && fnz != 0) {
        // subnormal: rare, so a divergent branch is the cheap form
        const int L = bitlen<NL>(u.m);
        shl<NL>(u.m, (uint32_t)(Fm::P - L));
        u.e -= Fm::P - L;
    }
}

template <int NL>
__device__ __forceinline__ void enc_qnan(uint32_t (&o)[NL])
{
    using Fm = F<NL>;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o[i] = 0;
    o[NL - 1] = (Fm::EMASK << Fm::TOPF) | (1u << (Fm::TOPF - 1));
}

template <int NL>
__device__ __forceinline__ void enc_inf(uint32_t (&o)[NL], int sign)
{
    using Fm = F<NL>;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o[i] = 0;
    o[NL - 1] = ((uint32_t)sign << 31) | (Fm::EMASK << Fm::TOPF);
}

template <int NL>
__device__ __forceinline__ void enc_zero(uint32_t (&o)[NL], int sign)
{
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o[i] = 0;
    o[NL - 1] = (uint32_t)sign << 31;
}

template <int NL>
__device__ __forceinline__ void enc_maxnorm(uint32_t (&o)[NL], int sign)
{
    using Fm = F<NL>;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o[i] = 0xFFFFFFFFu;
    o[NL - 1] = ((uint32_t)sign << 31) | ((Fm::EMASK - 1u) << Fm::TOPF) | Fm::FMASK;
}

// host/src/softfloat.c:120-134
__device__ __forceinline__ bool round_incr(int rnd, int sign, bool g, bool s, bool lsb)
{
    switch (rnd) {
    case RNE: return g && (s || lsb);
    case RTZ: return false;
    case RDN: return sign && (g || s);
    case RUP: return !sign && (g || s);
    default:  return g;
    }
}

// host/src/softfloat.c:141-149
__device__ __forceinline__ bool overflow_to_inf(int rnd, int sign)
{
    return rnd == RTZ ? false : rnd == RDN ? sign != 0 : rnd == RUP ? sign == 0 : true;
}

// k (NL limbs, < 2^P) += 1.  Returns true when it carried to exactly 2^P.
template <int NL>
__device__ __forceinline__ bool inc_sig(uint32_t (&k)[NL])
{
    using Fm = F<NL>;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        c += k[i];
        k[i] = (uint32_t)c;
        c >>= 32;
    }
    return (k[NL - 1] >> Fm::TOPF) & 2u;      // bit P set
}

// Round and pack (-1)^sign * K.g s * 2^(e_norm - P + 1): K the top P bits of
// the exact value (bit P-1 set), g the next bit, s whether anything below it
// is nonzero, e_norm the exponent of K's leading bit.  host/src/softfloat.c:
// 165-261, with tininess decided on the same rounding at an unbounded range.
template <int NL>
__device__ __forceinline__ void round_pack_norm(int sign, uint32_t (&K)[NL], bool g, bool s,
                                                int e_norm, int rnd, uint32_t (&out)[NL],
                                                uint32_t &flags)
{
    using Fm = F<NL>;
    if (e_norm >= Fm::EMIN) {
        const bool inx = g || s;
        if (round_incr(rnd, sign, g, s, K[0] & 1u)) {
            if (inc_sig<NL>(K)) {               // 2^P: renormalise, exact
#pragma unroll
                for (int i = 0; i < NL - 1; i++)
                    K[i] = 0;
                K[NL - 1] = 1u << Fm::TOPF;
                e_norm++;
            }
        }
        if (e_norm > Fm::EMAX) {
            if (overflow_to_inf(rnd, sign))
                enc_inf<NL>(out, sign);
            else
                enc_maxnorm<NL>(out, sign);
            flags |= FL_OF | FL_NX;
            return;
        }
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            out[i] = K[i];
        out[NL - 1] = ((uint32_t)sign << 31) | ((uint32_t)(e_norm + Fm::BIAS) << Fm::TOPF) |
                      (K[NL - 1] & Fm::FMASK);
        flags |= inx ? FL_NX : 0u;
        return;
    }
    // Below 2^emin before rounding: tiny unless rounding at full precision
    // carries exactly up to 2^emin.
    bool allones = (K[NL - 1] | ~((2u << Fm::TOPF) - 1u)) == 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        allones = allones && K[i] == 0xFFFFFFFFu;
    const bool carry_u = allones && round_incr(rnd, sign, g, s, K[0] & 1u);
    const bool tiny = !(e_norm == Fm::EMIN - 1 && carry_u);
    // Denormalise X = K.g by sh = emin - e_norm places, sticky folded.
    uint32_t X[NL + 1];
    X[0] = g ? 1u : 0u;                         // X = (K << 1) | g ... in two words
#pragma unroll
    for (int i = 0; i < NL; i++) {
        X[i] |= K[i] << 1;
        X[i + 1] = K[i] >> 31;
    }
    uint32_t jam = s ? 1u : 0u;
    shr_jam<NL + 1>(X, (uint32_t)(Fm::EMIN - e_norm), jam);
    const bool g2 = X[0] & 1u, s2 = jam != 0;
    uint32_t K2[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        K2[i] = alignbit(X[i + 1], X[i], 1);
    const bool inx = g2 || s2;
    if (round_incr(rnd, sign, g2, s2, K2[0] & 1u))
        inc_sig<NL>(K2);                         // may reach 2^(P-1): min normal
    // K2 < 2^P-1 ... 2^(P-1): packing K2 with a zero exponent field gives the
    // subnormal, or exponent field 1 when rounding reached 2^emin.
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        out[i] = K2[i];
    out[NL - 1] = ((uint32_t)sign << 31) | K2[NL - 1];
    if (inx)
        flags |= FL_NX | (tiny ? FL_UF : 0u);
    if (is_zero<NL>(K2))
        flags |= FL_UF;                          // rounded to zero (inexact)
}

// Round the exact nonzero value (-1)^sign * S * 2^eW, S an N-limb integer.
template <int NL, int N>
__device__ __forceinline__ void round_window(int sign, uint32_t (&S)[N], int eW, int rnd,
                                             uint32_t (&out)[NL], uint32_t &flags)
{
    using Fm = F<NL>;
    static_assert(N >= NL, "window narrower than the format");
    const int L = bitlen<N>(S);
    shl<N>(S, (uint32_t)(32 * N - L));           // leading bit to bit 32N-1
    // the top NL limbs hold 32NL = P + EXPW bits: K, then g, then EXPW-1
    // sticky bits; every lower limb is sticky too
    constexpr int D = Fm::EXPW;
    uint32_t K[NL];
    K[NL - 1] = S[N - 1] >> D;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        K[i] = alignbit(S[N - NL + i + 1], S[N - NL + i], D);
    const bool g = (S[N - NL] >> (D - 1)) & 1u;
    uint32_t sticky = S[N - NL] & ((1u << (D - 1)) - 1u);
#pragma unroll
    for (int i = 0; i < N - NL; i++)
        sticky |= S[i];
    round_pack_norm<NL>(sign, K, g, sticky != 0, eW + L - 1, rnd, out, flags);
}

// ---- fused multiply-add -----------------------------------------------------

// a*b + c for the steered operands of OP, one rounding.  host/src/softfloat.c:
// 324-428 (special-case order: NaN, inf*0, inf-inf, inf, c inf, a zero factor)
// and 438-468 (ADD = a*1 + c, SUB = a*1 - c, MUL = a*b + (+-0)).
//
// Method.  Subnormal operands are normalised first (a rare branch), so the
// exact product P = ma*mb has its leading bit at 2P-2 or 2P-1.  The exact sum
// is formed in a window of WL limbs: bit 0 is reserved for a jam (sticky)
// bit, P sits at bits [1, 2P+1) unshifted, and the addend C is placed by ONE
// variable right shift of C << (2P+3), every bit that falls to bit 0 or below
// jammed into bit 0.  When C leads P by more than P+2 binades P is below
// every rounding boundary of the result and is itself replaced by the jam bit
// (C then sits at its top placement).  Jamming is exact in effect: the
// representative lies in the same open interval between consecutive rounding
// boundaries, which are all multiples of 2 in window units.
template <int NL, int OP>
__device__ __forceinline__ void fma_op(const uint32_t (&xa)[NL], const uint32_t (&xb)[NL],
                                       const uint32_t (&xc)[NL], int rnd, uint32_t (&out)[NL],
                                       uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P, WL = Fm::WL;
    Unp<NL> ua, ub, uc;
    unpack<NL>(xa, ua);
    if (OP == OP_ADD || OP == OP_SUB) {          // b := 1.0
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            ub.m[i] = 0;
        ub.m[NL - 1] = 1u << Fm::TOPF;
        ub.e = -Fm::MANW;
        ub.sign = 0;
        ub.zero = ub.inf = ub.nan = ub.snan = false;
    } else {
        unpack<NL>(xb, ub);
    }
    if (OP == OP_MUL) {                          // c := +-0, sign of the product
#pragma unroll
        for (int i = 0; i < NL; i++)
            uc.m[i] = 0;
        uc.e = 0;
        uc.sign = ua.sign ^ ub.sign;
        uc.zero = true;
        uc.inf = uc.nan = uc.snan = false;
    } else {
        unpack<NL>(xc, uc);
        if (OP == OP_SUB)
            uc.sign ^= 1;
    }
    const int sp = ua.sign ^ ub.sign;

    if (ua.nan || ub.nan || uc.nan || ua.inf || ub.inf || uc.inf || ua.zero || ub.zero) {
        if (ua.nan || ub.nan || uc.nan) {
            enc_qnan<NL>(out);
            flags |= (ua.snan || ub.snan || uc.snan) ? FL_INV : 0u;
        } else if (ua.inf || ub.inf) {
            if (ua.zero || ub.zero || (uc.inf && uc.sign != sp)) {
                enc_qnan<NL>(out);
                flags |= FL_INV;
            } else {
                enc_inf<NL>(out, sp);
            }
        } else if (uc.inf) {
            enc_inf<NL>(out, uc.sign);
        } else if (uc.zero) {
            enc_zero<NL>(out, uc.sign == sp ? uc.sign : (rnd == RDN));
        } else {
            // 0*b + c is c, bits and all (the steered c: SUB's flipped sign)
#pragma unroll
            for (int i = 0; i < NL - 1; i++)
                out[i] = xc[i];
            out[NL - 1] = xc[NL - 1] ^ (OP == OP_SUB ? 0x80000000u : 0u);
        }
        return;
    }

    uint32_t Pm[2 * NL];
    if (OP == OP_ADD || OP == OP_SUB) {
        // ma * 2^(P-1): a constant shift instead of a product
#pragma unroll
        for (int i = 0; i < 2 * NL; i++)
            Pm[i] = 0;
        constexpr int LS = (P - 1) / 32, BS = (P - 1) % 32;
#pragma unroll
        for (int i = 0; i <= NL; i++) {
            const uint32_t lo = (i - 1 >= 0) ? ua.m[i - 1] : 0u;
            const uint32_t hi = (i < NL) ? ua.m[i] : 0u;
            if (i + LS < 2 * NL)
                Pm[i + LS] = BS ? ((hi << BS) | (lo >> (32 - BS))) : hi;
        }
    } else {
        mul_limbs<NL, NL>(ua.m, ub.m, Pm);
    }
    const int ep = ua.e + ub.e;

    if (OP == OP_MUL) {
        round_window<NL, 2 * NL>(sp, Pm, ep, rnd, out, flags);
        return;
    }

    // window
    constexpr int TB = 2 * P - 1;                // the product's possible top bit
    const int Lp = ((Pm[TB / 32] >> (TB % 32)) & 1u) ? 2 * P : 2 * P - 1;
    const int VEp = ep + Lp - 1;
    const bool czero = uc.zero;
    const int VEc = uc.e + P - 1;
    const bool cdom = !czero && (VEc - VEp > P + 2);
    uint32_t S[WL], Cw[WL];
#pragma unroll
    for (int i = 0; i < WL; i++)
        S[i] = 0;
    // P << 1, or the jam bit when C dominates
#pragma unroll
    for (int i = 0; i < 2 * NL; i++)
        S[i] = (Pm[i] << 1) | (i > 0 ? Pm[i - 1] >> 31 : 0u);
    if (2 * NL < WL)
        S[2 * NL] = Pm[2 * NL - 1] >> 31;
    if (cdom) {
#pragma unroll
        for (int i = 0; i < WL; i++)
            S[i] = 0;
        S[0] = 1;
    }
    // C << (2P+3), then right by r
    {
        constexpr int CP = 2 * P + 3, CL = CP / 32, CB = CP % 32;
#pragma unroll
        for (int i = 0; i < WL; i++)
            Cw[i] = 0;
#pragma unroll
        for (int i = 0; i <= NL; i++) {
            const uint32_t lo = (i - 1 >= 0) ? uc.m[i - 1] : 0u;
            const uint32_t hi = (i < NL) ? uc.m[i] : 0u;
            if (i + CL < WL)
                Cw[i + CL] = CB ? ((hi << CB) | (lo >> (32 - CB))) : hi;
        }
        const int rr = cdom ? 0 : (2 * P + 2 - (uc.e - ep));
        uint32_t jam = 0;
        shr_jam<WL>(Cw, czero ? 0u : (uint32_t)rr, jam);
        Cw[0] |= jam ? 1u : 0u;
    }
    const int eW = cdom ? (uc.e - (2 * P + 3)) : (ep - 1);
    int sign;
    if (uc.sign == sp) {
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < WL; i++) {
            c += (uint64_t)S[i] + Cw[i];
            S[i] = (uint32_t)c;
            c >>= 32;
        }
        sign = sp;
    } else {
        // D = C - P; negative: take P - C
        uint32_t D[WL];
        uint32_t br = 0;
#pragma unroll
        for (int i = 0; i < WL; i++) {
            const uint64_t t = (uint64_t)Cw[i] - S[i] - br;
            D[i] = (uint32_t)t;
            br = (uint32_t)(t >> 63);
        }
        if (br) {
            uint32_t b2 = 0;
#pragma unroll
            for (int i = 0; i < WL; i++) {
                const uint64_t t = (uint64_t)0 - D[i] - b2;
                S[i] = (uint32_t)t;
                b2 = (uint32_t)(t >> 63);
            }
            sign = sp;
        } else {
#pragma unroll
            for (int i = 0; i < WL; i++)
                S[i] = D[i];
            sign = uc.sign;
        }
        if (is_zero<WL>(S)) {                    // exact cancellation
            enc_zero<NL>(out, rnd == RDN);
            return;
        }
    }
    round_window<NL, WL>(sign, S, eW, rnd, out, flags);
}

} // namespace fx
