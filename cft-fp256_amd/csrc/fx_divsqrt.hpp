// Correctly-rounded division and square root on fixed limbs.
//
// The contract (python/cft_golden/softfloat.py:440-506, read as text; the
// reference library reaches the same bits through a Newton/Markstein
// composition, host/src/divsqrt.c): specials by operand class, otherwise the
// exact quotient / root carried past the rounding position with the
// remainder's nonzeroness as the sticky bit, then the one round_pack.  Any
// method that produces the exact floor and an exact "remainder != 0" gives
// the same bits and flags; the method here is ours:
//
//   division   schoolbook long division in radix 2^32 (Knuth's algorithm D):
//              each quotient digit estimated from the top limbs in binary64,
//              then an exact multiply-subtract and add-back, so the estimate
//              only decides how many corrections run.
//   sqrt       the radix-2^32 digit recurrence: bring down two limbs, choose
//              the largest digit d with (2rB + d) d <= R, exactly, starting
//              one above a binary64 estimate and stepping down.
#pragma once

#include "fx_arith.hpp"

namespace fx {

// Knuth D: Q = floor(X / D), X = U (2*NL limbs, U[2NL] = 0 appended), D
// normalised (top bit of D[NL-1] set), quotient NL digits.  Returns Q and
// whether the remainder is nonzero.
template <int NL>
__device__ __forceinline__ void long_div(uint32_t (&U)[2 * NL + 1], const uint32_t (&D)[NL],
                                         uint32_t (&Q)[NL], bool &rem_nz)
{
    double dtop = (double)D[NL - 1] * 4294967296.0;
    if constexpr (NL >= 2)
        dtop += (double)D[NL - 2];
    const double rinv = recip_mid(dtop);         // dtop in [2^63, 2^64)
#pragma unroll
    for (int j = NL - 1; j >= 0; j--) {
        // estimate from U[j+NL], U[j+NL-1], U[j+NL-2] against D's top two.
        // At or above the digit: num, dtop, rinv and the product each carry
        // one rounding, so num * rinv is within 5 * 2^-53 relative, 2^-18.7
        // absolute (the quotient is below 2^32 * (1 + 2^-31)), of U3/D2 (the
        // top three remainder limbs over the top two divisor limbs), and the
        // true digit is at most floor((U3 + 1)/D2), 2^-63 above U3/D2, and at
        // most B - 1.  So + 2^-17, capped at B - 1, never undershoots, and
        // overshoots only when U3/D2 lies within 2^-17 below an integer (or
        // by Knuth's truncated-divisor bound, also for the cap): the add-back
        // below almost never runs.  The estimate is never negative, so one
        // minimum is the whole clamp.
        const double num =
            __builtin_fma((double)U[j + NL - 1], 0x1p32, (double)U[j + NL] * 0x1p64) +
            (double)(j + NL - 2 >= 0 ? U[(j + NL - 2) < 0 ? 0 : j + NL - 2] : 0u);
        uint32_t qhat = (uint32_t)__builtin_fmin(num * rinv + 0x1p-17, 4294967295.0);
        // U[j .. j+NL] -= qhat * D: the NL partial products independently,
        // their limbs assembled by one carry chain and subtracted by one
        // borrow chain
        uint32_t pl[NL], ph[NL];
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const uint64_t p = (uint64_t)qhat * D[i];
            pl[i] = (uint32_t)p;
            ph[i] = (uint32_t)(p >> 32);
        }
        uint32_t c = 0, br = 0;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            uint32_t co, bo;
            const uint32_t pi = addc(pl[i], i ? ph[i - 1] : 0u, c, &co);
            c = co;
            U[j + i] = subc(U[j + i], pi, br, &bo);
            br = bo;
        }
        {
            uint32_t bo;                         // qhat * D < 2^(32(NL+1)): no carry past
            U[j + NL] = subc(U[j + NL], ph[NL - 1] + c, br, &bo);
            br = bo;
        }
        // negative: add D back until it is not (the exact test decides)
        bool neg = br != 0;
        while (neg) {
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i < NL; i++) {
                uint32_t co;
                U[j + i] = addc(U[j + i], D[i], c, &co);
                c = co;
            }
            uint32_t co;
            U[j + NL] = addc(U[j + NL], 0u, c, &co);
            neg = co == 0;                       // no carry out: still negative
            qhat--;
        }
        Q[j] = qhat;
    }
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        o |= U[i];
    rem_nz = o != 0;
}

// An operand the finite cores take as it is: finite, nonzero and normal (its
// exponent field neither 0 nor all ones).  Anything else goes through the
// full div_op / sqrt_op, which unpack, settle the specials and normalise.
template <int NL>
__device__ __forceinline__ bool plain_operand(const uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    const uint32_t ef = (x[NL - 1] >> Fm::TOPF) & Fm::EMASK;
    return ef != 0 && ef != Fm::EMASK;
}

// The quotient of finite nonzero operands, unpacked and normalised, as
// round_pack_norm takes it: K (P bits), guard, sticky and K's exponent.
template <int NL>
__device__ __forceinline__ void div_quotient(const Unp<NL> &ua, const Unp<NL> &ub,
                                             uint32_t (&K)[NL], bool &g, bool &s, int &e_norm)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P;
    // Q = floor(ma * 2^k / mb) with exactly P+3 bits: k = P+2 when ma >= mb,
    // which is no borrow out of ma - mb (one borrow chain, the differences
    // unused)
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        uint32_t bo;
        (void)subc(ua.m[i], ub.m[i], br, &bo);
        br = bo;
    }
    const bool ge = br == 0;
    const int k = ge ? P + 2 : P + 3;
    constexpr int SD = 32 * NL - P;              // divisor normalisation
    uint32_t D[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        D[i] = ub.m[i];
    shl<NL>(D, SD);
    // U = ma << (k + SD) = (ma << (k + SD - 32NL)) in the top NL limbs
    // k + SD - 32NL is 3, or 2 when ma >= mb: one funnel shift a limb (ma's
    // top bit, P-1, stays inside the top limb either way)
    static_assert(P + 3 <= 32 * NL, "the shifted dividend must fit its limbs");
    uint32_t U[2 * NL + 1];
#pragma unroll
    for (int i = 0; i < NL; i++) {
        U[i] = 0;
        U[NL + i] = alignbit(ua.m[i], i ? ua.m[i - 1] : 0u, ge ? 30u : 29u);
    }
    U[2 * NL] = 0;
    uint32_t Qd[NL];
    bool rem_nz;
    long_div<NL>(U, D, Qd, rem_nz);
    // Q has P+3 bits: K = Q >> 3, guard bit 2, sticky bits 1..0 and the rest
#pragma unroll
    for (int i = 0; i < NL; i++)
        K[i] = alignbit(i + 1 < NL ? Qd[i + 1] : 0u, Qd[i], 3);
    g = (Qd[0] >> 2) & 1u;
    s = (Qd[0] & 3u) != 0 || rem_nz;
    e_norm = ua.e - ub.e - k + P + 2;
}

// a / b for finite nonzero operands, unpacked and normalised.
template <int NL>
__device__ __forceinline__ void div_finite(const Unp<NL> &ua, const Unp<NL> &ub, int rnd,
                                           uint32_t (&out)[NL], uint32_t &flags)
{
    uint32_t K[NL];
    bool g, s;
    int e_norm;
    div_quotient<NL>(ua, ub, K, g, s, e_norm);
    round_pack_norm<NL>(ua.sign ^ ub.sign, K, g, s, e_norm, rnd, out, flags);
}

template <int NL>
__device__ __forceinline__ void div_op(const uint32_t (&xa)[NL], const uint32_t (&xb)[NL], int rnd,
                                       uint32_t (&out)[NL], uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P;
    Unp<NL> ua, ub;
    unpack_raw<NL>(xa, ua);
    unpack_raw<NL>(xb, ub);
    const int sq = ua.sign ^ ub.sign;
    if (ua.nan || ub.nan) {
        enc_qnan<NL>(out);
        flags |= (ua.snan || ub.snan) ? FL_INV : 0u;
        return;
    }
    if (ua.inf) {
        if (ub.inf) { enc_qnan<NL>(out); flags |= FL_INV; }
        else enc_inf<NL>(out, sq);
        return;
    }
    if (ub.inf) { enc_zero<NL>(out, sq); return; }
    if (ub.zero) {
        if (ua.zero) { enc_qnan<NL>(out); flags |= FL_INV; }
        else { enc_inf<NL>(out, sq); flags |= FL_DZ; }
        return;
    }
    if (ua.zero) { enc_zero<NL>(out, sq); return; }
    // A subnormal dividend (below 2^emin) over a normal divisor of 2^P or
    // more (lsb exponent 1 or more) is a quotient P+1 or more places below
    // 2^emin; a normal dividend of 4 or more (leading exponent 2 or more)
    // over a subnormal divisor is one of 2^(emax+1) or more.  Either rounds
    // by its exponent alone (to +-0 or the least subnormal, or to an infinity
    // or the largest finite number), so those lanes skip the normalisation
    // and the division and give round_pack_norm just an exponent that says
    // so (K is then read for nothing but a carry).
    const bool tiny = ua.sub && !ub.sub && ub.e >= 1;
    const bool huge = ub.sub && !ua.sub && ua.e + Fm::MANW >= 2;
    uint32_t K[NL];
    bool g = false, s = true;
    int e_norm;
    if (tiny || huge) {
#pragma unroll
        for (int i = 0; i < NL; i++)
            K[i] = ua.m[i];
        e_norm = tiny ? Fm::EMIN - P - 1 : Fm::EMAX + 1;
    } else {
        normalise<NL>(ua);
        normalise<NL>(ub);
        div_quotient<NL>(ua, ub, K, g, s, e_norm);
    }
    round_pack_norm<NL>(sq, K, g, s, e_norm, rnd, out, flags);
}

// r = floor(sqrt(M)), M 2*NL limbs with its top bit at 64NL-1 or 64NL-2;
// r has NL limbs with its top bit set.  rem_nz = (r*r != M).
template <int NL>
__device__ __forceinline__ void isqrt_limbs(const uint32_t (&M)[2 * NL], uint32_t (&r)[NL],
                                            bool &rem_nz)
{
    // R: the running remainder, at most NL + 2 limbs; r grows one digit a step
    uint32_t R[NL + 2];
#pragma unroll
    for (int i = 0; i < NL + 2; i++)
        R[i] = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        r[i] = 0;
    // first digit: isqrt of the top 64 bits.  g as a double (one rounding,
    // 2^-53 relative), its root (one more): within 2^-21 of sqrt(g) < 2^32,
    // so the truncated root is floor(sqrt(g)) or one off either way, and one
    // exact step each way settles it.
    {
        const uint64_t g = ((uint64_t)M[2 * NL - 1] << 32) | M[2 * NL - 2];
        const double gd = __builtin_fma((double)M[2 * NL - 1], 0x1p32, (double)M[2 * NL - 2]);
        uint32_t d = (uint32_t)__builtin_fmin(sqrt_mid(gd), 4294967295.0);   // gd >= 2^62
        uint64_t sq = (uint64_t)d * d;
        const bool over = sq > g;
        const uint64_t up = sq + 2ull * d + 1ull;        // (d+1)^2
        const bool under = !over && d != 0xFFFFFFFFu && up <= g;
        sq = over ? sq - 2ull * d + 1ull : under ? up : sq;
        d = over ? d - 1u : under ? d + 1u : d;
        const uint64_t rem = g - sq;
        R[0] = (uint32_t)rem;
        R[1] = (uint32_t)(rem >> 32);
        r[NL - 1] = d;                           // r kept left-aligned: digits fill down
    }
    // 1 / (V B) from the third digit on, scaled by 2^(32(k-1)) like Rs below:
    // set at the third digit and, so scaled, the same for every later one
    double rinv = 0.0;
#pragma unroll
    for (int k = 1; k < NL; k++) {
        // r so far: digits r[NL-k .. NL-1] (k limbs); R < 2^33 * 2^(32(k-1))
        // bring down two limbs: R = R * B^2 + (M[2NL-2k-1] : M[2NL-2k-2])
#pragma unroll
        for (int i = NL + 1; i >= 2; i--)
            R[i] = R[i - 2];
        R[1] = M[2 * NL - 2 * k - 1];
        R[0] = M[2 * NL - 2 * k - 2];
        // V = 2r as k+1 limbs (low first)
        uint32_t V[NL + 1];
#pragma unroll
        for (int i = 0; i <= NL; i++)
            V[i] = 0;
#pragma unroll
        for (int i = 0; i < k; i++) {
            const uint32_t lo = i > 0 ? r[NL - k + i - 1] : 0u;
            V[i] = alignbit(r[NL - k + i], lo, 31);
        }
        V[k] = r[NL - 1] >> 31;
        // estimate d = R / (V * B), from the top limbs as doubles.  Rs is R's
        // limbs k-1 .. k+2 summed low to high (each step one rounding) at
        // the scale 2^(-32(k-1)); Vd (the second and third digits) V B at
        // the scale 2^(-32(k-1)).  Powers of two commute with rounding, so
        // these are the unscaled sums' bits, scaled.
        double Rs = (double)R[k - 1];
        Rs = __builtin_fma((double)R[k], 0x1p32, Rs);
        Rs = __builtin_fma((double)R[k + 1], 0x1p64, Rs);
        Rs = __builtin_fma((double)R[k + 2], 0x1p96, Rs);
        // At or above the digit: R's dropped limbs move R / (V B) by under
        // 2^-63 (V B >= 2^64), V's only raise the estimate, and the digit is
        // floor(x) for the root x of (V B + x) x = R.
        // Second digit (V B = 2 r B, 33 significant bits, exact as a double,
        // rinv its correctly rounded reciprocal): with q = R / (V B) and
        // u = x / (V B) <= 2^-32, x = q / (1 + u) and q - q^2 / (V B) =
        // q (1 - u - u^2), so the one Newton step q - q * (q * rinv) lies at
        // most 2 q u^2 <= 2^-30 below x; q carries five roundings (Rs three,
        // rinv, the product), 2^-17.7 absolute at q < 2^33, the step two more
        // (2^-21): + 2^-15 never undershoots, and overshoots by less than 1,
        // so the add-back below runs only when x lies within 2^-15 of an
        // integer from below.
        // Third digit: V B there is (the second's V B + 2 d1) at the running
        // scale, so its reciprocal is rinv / (1 + e), e = 2 d1 rinv <
        // 2^-31, taken as rinv (1 - e): low by under e^2 = 2^-62 relative
        // (2^-29 absolute on the estimate) with two roundings.
        // From the fourth digit on, V B is the third digit's times
        // 2^(32(k-2)), its reciprocal the third's times 2^(-32(k-2)) (at the
        // running scale, the third's itself): that drops V's limbs below (each
        // digit's 2d, under 2^-63 of V B), which only raises the estimate, by
        // under 2^-30.  Rs * rinv carries at most 8 * 2^-53 relative of
        // rounding, 2^-17 absolute, which with the 2^-29 is below the + 2^-15:
        // it never undershoots.
        double dd;
        if (k == 1) {
            rinv = recip_mid(__builtin_fma((double)V[1], 0x1p64, (double)V[0] * 0x1p32));
            const double qd = Rs * rinv;
            dd = __builtin_fma(-qd, qd * rinv, qd) + 0x1p-15;
        } else {
            if (k == 2)
                rinv = __builtin_fma(-rinv, 2.0 * (double)r[NL - 2] * rinv, rinv);
            dd = Rs * rinv + 0x1p-15;
        }
        uint32_t d = (uint32_t)__builtin_fmin(dd, 4294967295.0);   // dd > 0
        // R -= T, T = (V*B + d) * d: the limbs of V*B + d are [d, V0 .. Vk];
        // their products with d are formed independently, assembled by one
        // add-with-carry chain and subtracted by one borrow chain.  R lies
        // below 2^(32(k+2)+1) (the last remainder, under 2r+1 < 2^(32k+1),
        // brought down two limbs) and T below 2^(32(k+3)), so limbs k+3 up
        // are zero before and after: the chains stop at limb k+2, its borrow
        // out standing for R's sign (R = R[0..k+2] - br 2^(32(k+3))).
        uint32_t pl[NL + 1], ph[NL + 1];
#pragma unroll
        for (int i = 0; i <= k + 1; i++) {
            const uint64_t p = (uint64_t)(i == 0 ? d : V[i - 1]) * d;
            pl[i] = (uint32_t)p;
            ph[i] = (uint32_t)(p >> 32);
        }
        uint32_t br = 0;
        {
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i <= k + 1; i++) {
                uint32_t co, bo;
                const uint32_t ti = addc(pl[i], i ? ph[i - 1] : 0u, c, &co);
                c = co;
                R[i] = subc(R[i], ti, br, &bo);
                br = bo;
            }
            uint32_t bo;
            R[k + 2] = subc(R[k + 2], ph[k + 1] + c, br, &bo);
            br = bo;
        }
        // while R < 0: R += V*B + 2d - 1 (= T(d) - T(d - 1)); d--.  A carry
        // out of limb k+2 cancels the borrow: R is then no longer negative.
        while (br) {
            uint32_t Sv[NL + 2];
            {
                const uint64_t low = 2ull * d - 1ull;
                Sv[0] = (uint32_t)low;
                uint64_t c = low >> 32;
#pragma unroll
                for (int i = 1; i <= k + 2; i++) {
                    c += (i - 1 <= k) ? V[i - 1] : 0u;
                    Sv[i] = (uint32_t)c;
                    c >>= 32;
                }
            }
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i <= k + 2; i++) {
                uint32_t co;
                R[i] = addc(R[i], Sv[i], c, &co);
                c = co;
            }
            br = c ^ 1u;                         // no carry out: still negative
            d--;
        }
        r[NL - 1 - k] = d;
    }
    // r is left-aligned in r[0..NL) (all NL digits placed)
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < NL + 2; i++)
        o |= R[i];
    rem_nz = o != 0;
}

// sqrt(a) for a positive normal operand, unpacked and normalised.
template <int NL>
__device__ __forceinline__ void sqrt_finite(const Unp<NL> &ua, int rnd, uint32_t (&out)[NL],
                                            uint32_t &flags)
{
    using Fm = F<NL>;
    constexpr int P = Fm::P;
    // M = m << t, top bit at 64NL-1 or 64NL-2, (e - t) even: t is the
    // constant 64NL - P, or one less; both shifts move m by the same whole
    // limbs Q and then R or R-1 bits, one funnel shift a limb
    constexpr int SH = 64 * NL - P, Q = SH / 32, R = SH % 32;
    static_assert(R >= 2, "both shifts within one limb");
    const uint32_t back = (uint32_t)((ua.e - SH) & 1);
    const int t = SH - (int)back;
    const uint32_t rb = (uint32_t)R - back;
    uint32_t M[2 * NL];
#pragma unroll
    for (int i = 0; i < 2 * NL; i++) {
        const int j = i - Q;
        const uint32_t hi = (j >= 0 && j < NL) ? ua.m[j] : 0u;
        const uint32_t lo = (j - 1 >= 0 && j - 1 < NL) ? ua.m[j - 1] : 0u;
        M[i] = (j < 0 || j > NL) ? 0u : alignbit(hi, lo, 32u - rb);
    }
    uint32_t r[NL];
    bool rem_nz;
    isqrt_limbs<NL>(M, r, rem_nz);
    // r in [2^(32NL-1), 2^(32NL)): K = top P bits, then guard and sticky
    constexpr int DROP = 32 * NL - P;            // = EXPW
    uint32_t K[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        K[i] = alignbit(i + 1 < NL ? r[i + 1] : 0u, r[i], DROP);
    const bool g = (r[0] >> (DROP - 1)) & 1u;
    const bool s = (r[0] & ((1u << (DROP - 1)) - 1u)) != 0 || rem_nz;
    round_pack_norm<NL>(0, K, g, s, 32 * NL - 1 + (ua.e - t) / 2, rnd, out, flags);
}

template <int NL>
__device__ __forceinline__ void sqrt_op(const uint32_t (&xa)[NL], int rnd, uint32_t (&out)[NL],
                                        uint32_t &flags)
{
    Unp<NL> ua;
    unpack<NL>(xa, ua);
    if (ua.nan) {
        enc_qnan<NL>(out);
        flags |= ua.snan ? FL_INV : 0u;
        return;
    }
    if (ua.zero || (ua.inf && !ua.sign)) {       // +-0 and +inf are themselves
#pragma unroll
        for (int i = 0; i < NL; i++)
            out[i] = xa[i];
        return;
    }
    if (ua.sign) {                               // negative, -inf included
        enc_qnan<NL>(out);
        flags |= FL_INV;
        return;
    }
    sqrt_finite<NL>(ua, rnd, out, flags);
}

} // namespace fx
