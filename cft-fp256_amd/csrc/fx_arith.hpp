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
    // FMA window (2NL limbs): the bit T where the larger term's leading bit
    // is placed.  Chosen per format so that the product's placement is a limb
    // move plus a 1..32-bit shift, the three possible rounding positions
    // T-P .. T-P+2 share a limb with the guard bit, bit 0 stays below every
    // real bit (the jam), and T+2 stays under the sign bit 64NL-1.
    static constexpr int T = NL == 1 ? 57 : NL == 2 ? 120 : NL == 4 ? 248 : 504;
    static_assert(T + 2 <= 64 * NL - 1, "no room above the window's top term");
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
        // a column's first product sets hi to its carry (the last column's
        // hi went into acc's top half below)
        const int i0 = k < NB ? 0 : k - NB + 1;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            const int j = k - i;
            if (j < 0 || j >= NB)
                continue;
            if (i == i0)
                mac_first(acc, hi, x[i], y[j]);
            else
                mac(acc, hi, x[i], y[j]);
        }
        p[k] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)hi << 32);
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

// x <<= 32*k limbs, zero fill; 0 <= k <= N (a select level per power of
// two up to N: a shift by N limbs empties the array).
template <int N>
__device__ __forceinline__ void limb_shl(uint32_t (&x)[N], uint32_t k)
{
#pragma unroll
    for (int t = 16; t >= 1; t >>= 1) {
        if (t > N)
            continue;
        const bool s = (k & (uint32_t)t) != 0;
#pragma unroll
        for (int i = N - 1; i >= 0; i--)
            x[i] = s ? (i >= t ? x[i - t] : 0u) : x[i];
    }
}

// x >>= 32*k limbs, the dropped limbs ORed into jam; 0 <= k < N (a select
// level per power of two below N).
template <int N>
__device__ __forceinline__ void limb_shr_jam(uint32_t (&x)[N], uint32_t k, uint32_t &jam)
{
#pragma unroll
    for (int t = 16; t >= 1; t >>= 1) {
        if (t >= N)
            continue;
        const bool s = (k & (uint32_t)t) != 0;
        uint32_t dropped = 0;
#pragma unroll
        for (int i = 0; i < t; i++)
            dropped |= x[i];
        jam |= s ? dropped : 0u;
#pragma unroll
        for (int i = 0; i < N; i++)
            x[i] = s ? (i + t < N ? x[i + t] : 0u) : x[i];
    }
}

// x >>= r bits (any r >= 0), every bit shifted out ORed into jam (nonzero
// iff something was dropped).  A shift past 32N-1 is taken as 32N-1: every
// caller's top bit (bit 32N-1) is clear, so that empties x all the same.
template <int N>
__device__ __forceinline__ void shr_jam(uint32_t (&x)[N], uint32_t r, uint32_t &jam)
{
    if (r > 32u * N - 1u)
        r = 32u * N - 1u;
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
    bool zero, inf, nan, snan, sub;
};

// host/src/softfloat.c:85-116.  A subnormal keeps its raw fraction here and
// sets `sub`; normalise() below moves its leading bit to P-1.
template <int NL>
__device__ __forceinline__ void unpack_raw(const uint32_t (&x)[NL], Unp<NL> &u)
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
    u.sub = ef == 0 && fnz != 0;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        u.m[i] = x[i];
    u.m[NL - 1] = (top & Fm::FMASK) | (ef ? (1u << Fm::TOPF) : 0u);
    u.e = (ef ? (int)ef : 1) - Fm::BIAS - Fm::MANW;
}

// A subnormal's significand shifted so its leading bit is at P-1, its
// exponent lowered to match; any other value is left alone.
//
// Whole limbs first: select levels of 4, 2 and 1 limbs, each taken while
// the value stays below 2^(P-1-32t), so that the leading bit ends 1 to 32
// places below P-1 (it starts at P-2 or lower, and NL-1 limbs bring even bit
// 0 that far).  Then one funnel shift by those v places, a limb at a time:
// v = 32 is the funnel's shift count 0, which yields the lower limb, the
// shift by a whole limb that v = 32 asks for.  No bit length and no second
// shift array, as a general left shift would need.
template <int NL>
__device__ __forceinline__ void normalise(Unp<NL> &u)
{
    using Fm = F<NL>;
    if (u.sub) {
        int k = 0;
#pragma unroll
        for (int t = 4; t >= 1; t >>= 1) {
            if (t > NL - 1)
                continue;
            // bits 32(NL-1-t) + TOPF and up all clear: below 2^(P-1-32t)
            uint32_t hi = u.m[NL - 1 - t] >> Fm::TOPF;
#pragma unroll
            for (int i = NL - t; i < NL; i++)
                hi |= u.m[i];
            const bool s = hi == 0;
#pragma unroll
            for (int i = NL - 1; i >= 0; i--)
                u.m[i] = s ? (i >= t ? u.m[i - t] : 0u) : u.m[i];
            k += s ? t : 0;
        }
        // the leading bit: limb NL-1 below TOPF, or limb NL-2 at TOPF or up
        int v;
        if constexpr (NL == 1) {
            v = Fm::TOPF - 31 + clz32(u.m[0]);
        } else {
            const uint32_t top = u.m[NL - 1];
            v = top ? Fm::TOPF - 31 + clz32(top) : Fm::TOPF + 1 + clz32(u.m[NL - 2]);
        }
        const uint32_t sh = 32u - (uint32_t)v;       // 0 when v = 32
#pragma unroll
        for (int i = NL - 1; i >= 0; i--)
            u.m[i] = alignbit(u.m[i], i ? u.m[i - 1] : 0u, sh);
        u.e -= 32 * k + v;
        u.sub = false;
    }
}

template <int NL>
__device__ __forceinline__ void unpack(const uint32_t (&x)[NL], Unp<NL> &u)
{
    unpack_raw<NL>(x, u);
    normalise<NL>(u);
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

// host/src/softfloat.c:120-134, as a truth table a mode over the index
// (sign, g, s, lsb): RNE g&(s|lsb) 0xE0E0, RTZ 0, RDN sign&(g|s) 0xFC00,
// RUP !sign&(g|s) 0x00FC, RMM g 0xF0F0.  A lookup, not a switch: where the
// mode differs by lane (the sequencer's records) a switch is a chain of
// divergent branches.
__device__ __forceinline__ bool round_incr(int rnd, int sign, bool g, bool s, bool lsb)
{
    const uint32_t idx = ((uint32_t)sign << 3) | ((uint32_t)g << 2) | ((uint32_t)s << 1) |
                         (uint32_t)lsb;
    const uint32_t t = rnd == RMM ? 0xF0F0u
                                  : (uint32_t)(0x00FCFC000000E0E0ull >> (16 * (rnd & 3)));
    return (t >> idx) & 1u;
}

// host/src/softfloat.c:141-149: bit 2*rnd + sign of 0x363 (RNE and RMM
// always, RDN for a negative result, RUP for a positive one, RTZ never).
__device__ __forceinline__ bool overflow_to_inf(int rnd, int sign)
{
    return (0x363u >> (2 * rnd + sign)) & 1u;
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
        {
            // K += the increment, one carry chain; a carry to 2^P leaves
            // every lower limb zero, so renormalising is the top limb
            uint32_t c = round_incr(rnd, sign, g, s, K[0] & 1u) ? 1u : 0u;
#pragma unroll
            for (int i = 0; i < NL; i++) {
                uint32_t co;
                K[i] = addc(K[i], 0u, c, &co);
                c = co;
            }
            const bool up = (K[NL - 1] >> Fm::TOPF) & 2u;
            K[NL - 1] = up ? (1u << Fm::TOPF) : K[NL - 1];
            e_norm += up ? 1 : 0;
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
    // P + 1 or more places below 2^emin: K.g shifts out whole, into the
    // sticky, so the result is +-0 or the least subnormal, tiny and inexact
    // (the general path below gives the same, through a shift that a wave
    // then skips unless one of its lanes lies within the subnormal range).
    if (Fm::EMIN - e_norm >= Fm::P + 1) {
#pragma unroll
        for (int i = 0; i < NL; i++)
            out[i] = 0;
        out[0] = round_incr(rnd, sign, false, true, false) ? 1u : 0u;
        out[NL - 1] |= (uint32_t)sign << 31;
        flags |= FL_NX | FL_UF;
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
    uint32_t X[NL + 1];                          // X = (K << 1) | g
    X[0] = (K[0] << 1) | (g ? 1u : 0u);
#pragma unroll
    for (int i = 1; i < NL; i++)
        X[i] = alignbit(K[i], K[i - 1], 31);
    X[NL] = K[NL - 1] >> 31;
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

// Carry chains: one v_add_co/v_addc (v_sub_co/v_subb) per limb.
template <int N>
__device__ __forceinline__ uint32_t add_limbs(uint32_t (&r)[N], const uint32_t (&x)[N],
                                              const uint32_t (&y)[N], uint32_t cin)
{
    uint32_t c = cin;
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint32_t co;
        r[i] = addc(x[i], y[i], c, &co);
        c = co;
    }
    return c;
}

// x = -x (two's complement), N limbs.
template <int N>
__device__ __forceinline__ void neg_limbs(uint32_t (&x)[N])
{
    uint32_t c = 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint32_t co;
        x[i] = addc(~x[i], 0u, c, &co);
        c = co;
    }
}

// Round the exact value held in S (N limbs) whose leading bit is at bit M,
// M in {T-1, T, T+1} for a T fixed at compile time, e_norm the exponent of
// that leading bit.  The three possible rounding positions M-P+1 share one
// limb, so K is one variable v_alignbit per limb.
template <int NL, int N, int T>
__device__ __forceinline__ void round_fixed(int sign, const uint32_t (&S)[N], int M, int e_norm,
                                            int rnd, uint32_t (&out)[NL], uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P;
    constexpr int LO = T - P;                    // lowest K position (M = T-1)
    constexpr int LI = LO / 32;
    static_assert((T - P + 2) / 32 == LI, "rounding positions straddle a limb");
    static_assert(LO % 32 >= 1, "guard bit must share the limb");
    static_assert(LI + NL < N + 1, "window too narrow");
    const uint32_t off = (uint32_t)(M - P + 1 - 32 * LI);   // 1..31
    uint32_t K[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        K[i] = alignbit(LI + i + 1 < N ? S[LI + i + 1] : 0u, S[LI + i], off);
    K[NL - 1] &= (2u << Fm::TOPF) - 1u;          // the P bits, nothing above
    const bool g = (S[LI] >> (off - 1)) & 1u;
    uint32_t sticky = S[LI] & ((1u << (off - 1)) - 1u);
#pragma unroll
    for (int i = 0; i < LI; i++)
        sticky |= S[i];
    round_pack_norm<NL>(sign, K, g, sticky != 0, e_norm, rnd, out, flags);
}

// a*b + c for the steered operands of op, one rounding.  host/src/softfloat.c:
// 324-428 (special-case order: NaN, inf*0, inf-inf, inf, c inf, a zero factor)
// and 438-468 (ADD = a*1 + c, SUB = a*1 - c, MUL = a*b + (+-0)).  op is
// uniform across the wave (a kernel constant, or an interpreted instruction).
//
// Method.  For normal factors the exact product P = ma*mb has its leading
// bit at 2P-2 or 2P-1.  The exact sum is formed in a window of 2NL limbs in
// which the LARGER term - by the exponent of bit 2P-2 of the product and of
// bit P-1 of c - sits with that bit at a fixed bit T = 64NL-7, and the
// smaller term is placed by one variable right shift, every bit shifted past
// the bottom jammed into bit 0 (which no real bit of either term occupies).
// Jamming is exact in effect: the representative lies in the same open
// interval between consecutive rounding boundaries.  The sum's leading bit is
// then at T-1, T or T+1 unless the terms were within a binade of each other
// and of opposite sign (deep cancellation) or the larger term holds a
// subnormal significand (its leading bit below T), which take a rare branch
// that normalises the sum by a full shift.  (MUL normalises its factors
// instead: it rounds from the product's top bit.)
//
// ELEM: the elementwise kernels' variant.  There nearly every wavefront holds
// special and subnormal lanes, so a subnormal operand whose term can be the
// larger is normalised up front (N limbs, not the sum's 2N) and the rare and
// common sums share one rounding.  The sequencer's kernels, whose wavefronts
// rarely hold a subnormal, keep the leaner control flow: the rare sum has a
// rounding of its own and subnormal terms go to the sum's normalisation,
// which their kernels hold in fewer registers and a lone wavefront runs in
// fewer instructions.
template <int NL, bool ELEM = false>
__device__ __forceinline__ void fma_core(int op, const uint32_t (&xa)[NL], const uint32_t (&xb)[NL],
                                         const uint32_t (&xc)[NL], int rnd, uint32_t (&out)[NL],
                                         uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P, WN = 2 * NL, T = Fm::T;
    Unp<NL> ua, ub, uc;
    unpack_raw<NL>(xa, ua);
    if (op == OP_ADD || op == OP_SUB) {          // b := 1.0
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            ub.m[i] = 0;
        ub.m[NL - 1] = 1u << Fm::TOPF;
        ub.e = -Fm::MANW;
        ub.sign = 0;
        ub.zero = ub.inf = ub.nan = ub.snan = ub.sub = false;
    } else {
        unpack_raw<NL>(xb, ub);
    }
    if (op == OP_MUL) {                          // c := +-0, sign of the product
#pragma unroll
        for (int i = 0; i < NL; i++)
            uc.m[i] = 0;
        uc.e = 0;
        uc.sign = ua.sign ^ ub.sign;
        uc.zero = true;
        uc.inf = uc.nan = uc.snan = uc.sub = false;
    } else {
        unpack_raw<NL>(xc, uc);
        if (op == OP_SUB)
            uc.sign ^= 1;
    }
    const int sp = ua.sign ^ ub.sign;

    if (ua.nan || ub.nan || uc.nan || ua.inf || ub.inf || uc.inf || ua.zero || ub.zero) {
        // the special result by selects, in the same order: a NaN operand
        // gives qNaN (INVALID for a signalling one); an infinite factor gives
        // qNaN and INVALID times zero or against an opposite infinite c, else
        // an infinity; then an infinite c; a zero product gives the zero of
        // the sum's sign rule when c is zero, else c itself, bits and all
        // (the steered c: SUB's flipped sign)
        const bool anynan = ua.nan || ub.nan || uc.nan;
        const bool pinf = ua.inf || ub.inf;
        const bool badprod = pinf && (ua.zero || ub.zero || (uc.inf && uc.sign != sp));
        const bool qnan = anynan || badprod;
        const bool inf = !qnan && (pinf || uc.inf);
        const bool zero = !qnan && !inf && uc.zero;
        const bool isc = !qnan && !inf && !zero;
        const uint32_t zsign = uc.sign == sp ? (uint32_t)uc.sign : (rnd == RDN ? 1u : 0u);
        const uint32_t top_inf = ((uint32_t)(pinf ? sp : uc.sign) << 31) | (Fm::EMASK << Fm::TOPF);
        const uint32_t top_nan = (Fm::EMASK << Fm::TOPF) | (1u << (Fm::TOPF - 1));
        const uint32_t top_c = xc[NL - 1] ^ (op == OP_SUB ? 0x80000000u : 0u);
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            out[i] = isc ? xc[i] : 0u;
        out[NL - 1] = qnan ? top_nan : inf ? top_inf : zero ? (zsign << 31) : top_c;
        flags |= (anynan ? (ua.snan || ub.snan || uc.snan) : badprod) ? FL_INV : 0u;
        return;
    }
    // MUL rounds from the product's top bit, which needs normal factors.  FMA,
    // ADD and SUB may leave a subnormal significand raw: each term is placed
    // by its lsb exponent, so a subnormal one merely has its leading bit
    // below T, and the sum then takes the general normalisation below.  That
    // renormalises all 2NL limbs of the sum, so in the elementwise variant a
    // subnormal operand whose term can be the larger one (by the exponents of
    // bit P-1 of c and bit 2P-2 of the product, within two binades) is
    // normalised first, NL limbs: the sum's rare branch is then left to
    // cancellations.
    if (op == OP_MUL) {
        // unless even the product's bit 2P-1 lies P+1 or more places below
        // 2^emin: then it rounds to +-0 or the least subnormal whatever its
        // bits, and round_pack_norm takes that from the exponent alone
        if ((ua.sub || ub.sub) && ua.e + ub.e + 2 * P - 1 > Fm::EMIN - P - 1) {
            normalise<NL>(ua);
            normalise<NL>(ub);
        }
    } else if (ELEM && (ua.sub || ub.sub || uc.sub)) {
        const int vp = ua.e + ub.e + 2 * P - 2, vc = uc.e + P - 1;
        const bool pmay = !uc.zero ? vp + 2 > vc : true, cmay = !uc.zero && vc + 2 > vp;
        if (ua.sub && pmay)
            normalise<NL>(ua);
        if (ub.sub && pmay)
            normalise<NL>(ub);
        if (uc.sub && cmay)
            normalise<NL>(uc);
    }

    // the exact product, leading bit at 2P-2 or 2P-1 for normal factors
    uint32_t Pm[WN];
    if (op == OP_ADD || op == OP_SUB) {
#pragma unroll
        for (int i = 0; i < WN; i++)
            Pm[i] = 0;
        constexpr int LS = (P - 1) / 32, BS = (P - 1) % 32;
#pragma unroll
        for (int i = 0; i <= NL; i++) {
            const uint32_t lo = (i - 1 >= 0) ? ua.m[i - 1] : 0u;
            const uint32_t hi = (i < NL) ? ua.m[i] : 0u;
            if (i + LS < WN)
                Pm[i + LS] = BS ? alignbit(hi, lo, 32u - BS) : hi;
        }
    } else {
        mul_limbs<NL, NL>(ua.m, ub.m, Pm);
    }
    const int ep = ua.e + ub.e;
    constexpr int TB = 2 * P - 1;                // the product's possible top bit
    const bool ptop = (Pm[TB / 32] >> (TB % 32)) & 1u;

    if (op == OP_MUL) {
        // leading bit at 2P-1 or 2P-2: round from that fixed position
        round_fixed<NL, WN, 2 * P - 2>(sp, Pm, ptop ? 2 * P - 1 : 2 * P - 2,
                                       ep + (ptop ? 2 * P - 1 : 2 * P - 2), rnd, out, flags);
        return;
    }

    // Pw = P with its leading bit at T (bit 2P-2 at T when the top bit is
    // clear, normal or not): a shift of T-TB or T-TB+1 bits
    uint32_t Pw[WN], Cw[WN];
    {
        constexpr int SH = T - TB;               // when ptop
        static_assert(SH >= 1, "no room for the jam bit");
        const uint32_t s = ptop ? (uint32_t)SH : (uint32_t)(SH + 1);
        // constant limb part and a 0/1-bit remainder: (SH, SH+1) = 32q + (r, r+1)
        constexpr int Q = SH / 32, R = SH % 32;
        static_assert(R >= 1 && R + 1 <= 32, "v_alignbit takes a shift of 0..31");
        const uint32_t r = s - 32u * Q;          // R or R+1, >= 1
#pragma unroll
        for (int i = 0; i < WN; i++) {
            const int j = i - Q;
            const uint32_t hi = (j >= 0 && j < WN) ? Pm[j] : 0u;
            const uint32_t lo = (j - 1 >= 0 && j - 1 < WN) ? Pm[j - 1] : 0u;
            Pw[i] = alignbit(hi, lo, 32u - r);
        }
    }
    const int VEp = ep + (ptop ? 2 * P - 1 : 2 * P - 2);
    // Cw = C with bit P-1 at T (a constant shift): its leading bit, or above
    // it for a subnormal C
    {
        constexpr int SH = T - (P - 1), Q = SH / 32, R = SH % 32;
#pragma unroll
        for (int i = 0; i < WN; i++) {
            const int j = i - Q;
            const uint32_t hi = (j >= 0 && j < NL) ? uc.m[j] : 0u;
            const uint32_t lo = (j - 1 >= 0 && j - 1 < NL) ? uc.m[j - 1] : 0u;
            Cw[i] = R ? alignbit(hi, lo, 32u - R) : hi;
        }
    }
    // a zero addend is below everything: the product alone, exactly
    const int VEc = uc.zero ? VEp - 4 * WN * 32 : uc.e + P - 1;
    const int d = VEc - VEp;
    const bool cbig = d > 0;
    uint32_t big[WN], small[WN];
#pragma unroll
    for (int i = 0; i < WN; i++) {
        big[i] = cbig ? Cw[i] : Pw[i];
        small[i] = cbig ? Pw[i] : Cw[i];
    }
    const int bsign = cbig ? uc.sign : sp;
    const bool subtract = uc.sign != sp;
    {
        uint32_t jam = 0;
        const int ad = d < 0 ? -d : d;
        shr_jam<WN>(small, (uint32_t)(ad > 32 * WN - 1 ? 32 * WN - 1 : ad), jam);
        small[0] |= jam ? 1u : 0u;
    }
    const int VEb = cbig ? VEc : VEp;            // the exponent of bit T
    uint32_t S[WN];
    {
        const uint32_t m = subtract ? 0xFFFFFFFFu : 0u;
#pragma unroll
        for (int i = 0; i < WN; i++)
            small[i] ^= m;
        add_limbs<WN>(S, big, small, subtract ? 1u : 0u);
    }
    // leading bit: T+1 (carry), T, T-1 (borrow) - or lower / negative
    const bool t1 = (S[(T + 1) / 32] >> ((T + 1) % 32)) & 1u;
    const bool t0 = (S[T / 32] >> (T % 32)) & 1u;
    const bool tm = (S[(T - 1) / 32] >> ((T - 1) % 32)) & 1u;
    const bool neg = (S[WN - 1] >> 31) & 1u;     // only when d == 0, subtracting
    int sign = bsign;
    int M = t1 ? T + 1 : (t0 ? T : T - 1);
    int e_norm = VEb + (M - T);
    if (neg || (!t1 && !t0 && !tm)) {
        // rare: equal leading exponents cancelled, the smaller term won, or
        // the larger term was subnormal (a raw significand, or a product
        // with a subnormal factor) and sits below T
        if (neg) {
            neg_limbs<WN>(S);
            sign ^= 1;
        }
        const int L = bitlen<WN>(S);
        if (L == 0) {                            // exact cancellation
            enc_zero<NL>(out, rnd == RDN);
            return;
        }
        shl<WN>(S, (uint32_t)(T + 1 - (L - 1)));  // leading bit to T+1
        M = T + 1;
        e_norm = VEb + (L - 1) - T;
        if constexpr (!ELEM) {
            round_fixed<NL, WN, T>(sign, S, M, e_norm, rnd, out, flags);
            return;
        }
    }
    // Whenever the jam bit was set, the sum's rounding position lies at bit 2
    // or above: the leading bit is at T-1 or higher with both terms normal;
    // when the smaller term was shifted out past bit 0 under a subnormal
    // larger one, the result is below 2^emin and rounds at C's bit 0, at
    // T-P+1; and a product with a subnormal factor that is the larger term
    // only drops bits of a C at least 2^(T-P+1) times smaller.
    round_fixed<NL, WN, T>(sign, S, M, e_norm, rnd, out, flags);
}

// a + c, or a - c when neg_c: the value, bits and flags of fma_core's ADD /
// SUB (a*1 + c, one rounding; host/src/softfloat.c:438-468 and the special
// order of 324-428), in an NL-limb window instead of the FMA's 2NL.
//
// Method.  Each significand (implicit bit for a normal, raw for a
// subnormal: no normalisation is needed to add) goes 3 bits up, so bits
// 0..2 lie below both terms; the term with the smaller lsb exponent is
// shifted right by the difference, every bit that falls off ORed into bit 0
// with the bit already there.  P + 4 <= 32 NL for every format, so the sum
// and its sign fit.  Unless the terms cancel deeply (lsb exponents within
// one, opposite signs) or both are tiny, the sum's leading bit is at P+1,
// P+2 or P+3: the guard is then at bit 1 or above, every rounding boundary a
// multiple of 2, and the jammed sum lies in the same open interval between
// consecutive boundaries as the exact one.  The deep cases shift nothing
// off, so they are exact and take the general normalisation.
template <int NL>
__device__ __forceinline__ void add_core(const uint32_t (&xa)[NL], const uint32_t (&xc)[NL],
                                         bool neg_c, int rnd, uint32_t (&out)[NL],
                                         uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P;
    static_assert(P + 4 <= 32 * NL, "the shifted sum must fit the window");
    Unp<NL> ua, uc;
    unpack_raw<NL>(xa, ua);
    unpack_raw<NL>(xc, uc);
    if (neg_c)
        uc.sign ^= 1;
    if (ua.nan || uc.nan || ua.inf || uc.inf || ua.zero) {
        if (ua.nan || uc.nan) {
            enc_qnan<NL>(out);
            flags |= (ua.snan || uc.snan) ? FL_INV : 0u;
        } else if (ua.inf) {
            if (uc.inf && uc.sign != ua.sign) {
                enc_qnan<NL>(out);
                flags |= FL_INV;
            } else {
                enc_inf<NL>(out, ua.sign);
            }
        } else if (uc.inf) {
            enc_inf<NL>(out, uc.sign);
        } else if (uc.zero) {
            enc_zero<NL>(out, uc.sign == ua.sign ? uc.sign : (rnd == RDN));
        } else {
            // 0 + c is c, bits and all (SUB's flipped sign)
#pragma unroll
            for (int i = 0; i < NL - 1; i++)
                out[i] = xc[i];
            out[NL - 1] = xc[NL - 1] ^ (neg_c ? 0x80000000u : 0u);
        }
        return;
    }
    // both finite, a nonzero; a zero c sits below a and adds nothing
    const bool cbig = uc.e > ua.e;
    const int d = cbig ? uc.e - ua.e : ua.e - uc.e;
    const int eX = cbig ? uc.e : ua.e;
    uint32_t X[NL], Y[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const uint32_t a3 = alignbit(ua.m[i], i ? ua.m[i - 1] : 0u, 29);   // m << 3
        const uint32_t c3 = alignbit(uc.m[i], i ? uc.m[i - 1] : 0u, 29);
        X[i] = cbig ? c3 : a3;
        Y[i] = cbig ? a3 : c3;
    }
    int sign = cbig ? uc.sign : ua.sign;
    const bool subtract = ua.sign != uc.sign;
    {
        uint32_t jam = 0;
        shr_jam<NL>(Y, (uint32_t)(d > 32 * NL ? 32 * NL : d), jam);
        Y[0] |= jam ? 1u : 0u;
    }
    uint32_t S[NL];
    {
        const uint32_t m = subtract ? 0xFFFFFFFFu : 0u;
#pragma unroll
        for (int i = 0; i < NL; i++)
            Y[i] ^= m;
        add_limbs<NL>(S, X, Y, subtract ? 1u : 0u);
    }
    constexpr int T1 = P + 1, T2 = P + 2, T3 = P + 3;
    const bool b3 = (S[T3 / 32] >> (T3 % 32)) & 1u;
    const bool b2 = (S[T2 / 32] >> (T2 % 32)) & 1u;
    const bool b1 = (S[T1 / 32] >> (T1 % 32)) & 1u;
    const bool neg = (S[NL - 1] >> 31) & 1u;     // only with equal lsb exponents
    if (neg || (!b3 && !b2 && !b1)) {
        // rare: deep cancellation or a tiny sum, exact either way
        if (neg) {
            neg_limbs<NL>(S);
            sign ^= 1;
        }
        if (is_zero<NL>(S)) {                    // exact cancellation
            enc_zero<NL>(out, rnd == RDN);
            return;
        }
        round_window<NL, NL>(sign, S, eX - 3, rnd, out, flags);
        return;
    }
    const int lead = b3 ? T3 : (b2 ? T2 : T1);
    const uint32_t off = (uint32_t)(lead - P + 1);          // 2, 3 or 4
    uint32_t K[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        K[i] = alignbit(i + 1 < NL ? S[i + 1] : 0u, S[i], off);
    K[NL - 1] &= (2u << Fm::TOPF) - 1u;
    const bool g = (S[0] >> (off - 1)) & 1u;
    const bool st = (S[0] & ((1u << (off - 1)) - 1u)) != 0;
    round_pack_norm<NL>(sign, K, g, st, eX - 3 + lead, rnd, out, flags);
}

template <int NL, int OP>
__device__ __forceinline__ void fma_op(const uint32_t (&xa)[NL], const uint32_t (&xb)[NL],
                                       const uint32_t (&xc)[NL], int rnd, uint32_t (&out)[NL],
                                       uint32_t &flags)
{
    fma_core<NL, true>(OP, xa, xb, xc, rnd, out, flags);
}

} // namespace fx
