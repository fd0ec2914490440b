// The non-arithmetic elementwise opcodes on fixed limbs: sign operations,
// 754-2019 minimum/maximum, quiet comparisons, select, the integer group and
// the divide/sqrt seeds.  host/src/softfloat.c:470-787 and 939-1009 define
// every bit and flag; unassigned opcodes answer the canonical quiet NaN with
// INVALID (softfloat.c:1000-1007).
#pragma once

#include "fx_arith.hpp"

namespace fx {

// ---- seed tables, built at compile time from their definitions ----------
//
// recip: the nearest integer to 2^28 / (2^10 + 2i + 1), ties to even, for
// i in [0, 512).  rsqrt: for j = odd*512 + i, the nearest integer to
// sqrt(2^44 / m2) with m2 = (2^10 + 2i + 1) << odd, rounded up exactly when
// (q + 1/2)^2 * m2 < 2^44 (host/src/softfloat.c:663-705 state the same
// values).

struct SeedTables {
    uint32_t recip[512];
    uint32_t rsqrt[1024];
};

constexpr SeedTables make_seed_tables()
{
    SeedTables t{};
    for (uint32_t i = 0; i < 512; i++) {
        const uint64_t den = 1024u + 2u * i + 1u, num = 1ull << 28;
        uint64_t q = num / den;
        const uint64_t r = num % den;
        if (2 * r > den || (2 * r == den && (q & 1)))
            q++;
        t.recip[i] = (uint32_t)q;
    }
    for (uint32_t j = 0; j < 1024; j++) {
        const uint32_t odd = j >> 9, i = j & 511u;
        const uint64_t m2 = (uint64_t)(1024u + 2u * i + 1u) << odd;
        uint64_t lo = 0, hi = 1u << 18;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (mid * mid * m2 <= (1ull << 44))
                lo = mid;
            else
                hi = mid;
        }
        if ((2 * lo + 1) * (2 * lo + 1) * m2 < (1ull << 46))
            lo++;
        t.rsqrt[j] = (uint32_t)lo;
    }
    return t;
}

__constant__ static const SeedTables kSeeds = make_seed_tables();

// ---- helpers -----------------------------------------------------------

template <int NL>
__device__ __forceinline__ bool enc_is_nan(const uint32_t (&x)[NL], bool &snan)
{
    using Fm = F<NL>;
    const uint32_t top = x[NL - 1];
    uint32_t fnz = top & Fm::FMASK;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        fnz |= x[i];
    const bool nan = ((top >> Fm::TOPF) & Fm::EMASK) == Fm::EMASK && fnz != 0;
    snan = nan && !((top >> (Fm::TOPF - 1)) & 1u);
    return nan;
}

template <int NL>
__device__ __forceinline__ bool mag_zero(const uint32_t (&x)[NL])
{
    uint32_t o = x[NL - 1] & 0x7FFFFFFFu;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o |= x[i];
    return o == 0;
}

// unsigned compare of the full words (sign bit included): -1, 0, 1
template <int NL>
__device__ __forceinline__ int ucmp(const uint32_t (&x)[NL], const uint32_t (&y)[NL])
{
    int r = 0;
#pragma unroll
    for (int i = 0; i < NL; i++)
        r = x[i] != y[i] ? (x[i] < y[i] ? -1 : 1) : r;
    return r;
}

// x < y for non-NaN encodings, -0 == +0 (host/src/softfloat.c:487-503)
template <int NL>
__device__ __forceinline__ bool numeric_lt(const uint32_t (&x)[NL], const uint32_t (&y)[NL])
{
    const int sx = x[NL - 1] >> 31, sy = y[NL - 1] >> 31;
    if (mag_zero<NL>(x) && mag_zero<NL>(y))
        return false;
    if (sx != sy)
        return sx != 0;
    uint32_t mx[NL], my[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
        mx[i] = x[i];
        my[i] = y[i];
    }
    mx[NL - 1] &= 0x7FFFFFFFu;
    my[NL - 1] &= 0x7FFFFFFFu;
    const int c = ucmp<NL>(mx, my);
    return sx ? c > 0 : c < 0;
}

template <int NL>
__device__ __forceinline__ void copy(uint32_t (&o)[NL], const uint32_t (&x)[NL])
{
#pragma unroll
    for (int i = 0; i < NL; i++)
        o[i] = x[i];
}

template <int NL>
__device__ __forceinline__ void enc_one(uint32_t (&o)[NL])
{
    using Fm = F<NL>;
#pragma unroll
    for (int i = 0; i < NL - 1; i++)
        o[i] = 0;
    o[NL - 1] = (uint32_t)Fm::BIAS << Fm::TOPF;
}

// (sign, r, e): r an nbits-bit integer with its top bit set, packed with
// the RNE rounding of round_pack and its flags discarded.
template <int NL>
__device__ __forceinline__ void seed_pack(int sign, uint32_t r, int nbits, int e,
                                          uint32_t (&o)[NL])
{
    using Fm = F<NL>;
    uint32_t K[NL];
#pragma unroll
    for (int i = 0; i < NL; i++)
        K[i] = 0;
    K[0] = r;
    shl<NL>(K, (uint32_t)(Fm::P - nbits));
    uint32_t dummy = 0;
    round_pack_norm<NL>(sign, K, false, false, e + nbits - 1, RNE, o, dummy);
}

template <int NL>
__device__ __forceinline__ uint32_t top9(const uint32_t (&x)[NL])
{
    using Fm = F<NL>;
    // fraction bits [MANW-9, MANW)
    constexpr int lo = Fm::MANW - 9;
    constexpr int li = lo / 32, bi = lo % 32;
    const uint32_t w0 = x[li];
    uint32_t w1 = 0;
    if constexpr (li + 1 < NL)
        w1 = x[li + 1];
    return alignbit(w1, w0, bi) & 511u;
}

// ---- the opcodes ---------------------------------------------------------

template <int NL>
__device__ __forceinline__ void minmax(const uint32_t (&a)[NL], const uint32_t (&b)[NL],
                                       bool want_max, bool number, uint32_t (&o)[NL],
                                       uint32_t &flags)
{
    bool sa, sb;
    const bool an = enc_is_nan<NL>(a, sa), bn = enc_is_nan<NL>(b, sb);
    if ((an && sa) || (bn && sb))
        flags |= FL_INV;
    if (an || bn) {
        if (!number || (an && bn))
            enc_qnan<NL>(o);
        else
            copy<NL>(o, an ? b : a);
        return;
    }
    if (mag_zero<NL>(a) && mag_zero<NL>(b)) {
        const int za = a[NL - 1] >> 31, zb = b[NL - 1] >> 31;
        if (za == zb)
            copy<NL>(o, a);
        else
            copy<NL>(o, (want_max ? !za : za) ? a : b);
        return;
    }
    const bool lt = numeric_lt<NL>(a, b);
    copy<NL>(o, want_max ? (lt ? b : a) : (lt ? a : b));
}

template <int NL>
__device__ __forceinline__ void misc_op(int op, const uint32_t (&a)[NL], const uint32_t (&b)[NL],
                                        const uint32_t (&c)[NL], uint32_t (&o)[NL],
                                        uint32_t &flags)
{
    using Fm = F<NL>;
    switch (op) {
    case 4:                                          // abs
        copy<NL>(o, a);
        o[NL - 1] &= 0x7FFFFFFFu;
        return;
    case 5:                                          // neg
        copy<NL>(o, a);
        o[NL - 1] ^= 0x80000000u;
        return;
    case 6:                                          // copysign
        copy<NL>(o, a);
        o[NL - 1] = (a[NL - 1] & 0x7FFFFFFFu) | (b[NL - 1] & 0x80000000u);
        return;
    case 7: minmax<NL>(a, b, false, false, o, flags); return;
    case 8: minmax<NL>(a, b, true, false, o, flags); return;
    case 9: minmax<NL>(a, b, false, true, o, flags); return;
    case 10: minmax<NL>(a, b, true, true, o, flags); return;
    case 11:                                         // select: c != 0 ? a : b
        copy<NL>(o, mag_zero<NL>(c) ? b : a);
        return;
    case 12: case 13: case 14: {                     // cmplt, cmple, cmpeq
        bool sa, sb;
        const bool an = enc_is_nan<NL>(a, sa), bn = enc_is_nan<NL>(b, sb);
        if ((an && sa) || (bn && sb))
            flags |= FL_INV;
        bool t = false;
        if (!an && !bn) {
            const bool lt = numeric_lt<NL>(a, b), gt = numeric_lt<NL>(b, a);
            t = op == 12 ? lt : op == 13 ? (lt || !gt) : (!lt && !gt);
        }
        if (t)
            enc_one<NL>(o);
        else
            enc_zero<NL>(o, 0);
        return;
    }
    case 16: case 17: case 18:
#pragma unroll
        for (int i = 0; i < NL; i++)
            o[i] = op == 16 ? (a[i] & b[i]) : op == 17 ? (a[i] | b[i]) : (a[i] ^ b[i]);
        return;
    case 19: {                                       // iadd mod 2^width
        uint64_t cy = 0;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            cy += (uint64_t)a[i] + b[i];
            o[i] = (uint32_t)cy;
            cy >>= 32;
        }
        return;
    }
    case 20: {                                       // isub mod 2^width
        uint32_t br = 0;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const uint64_t t = (uint64_t)a[i] - b[i] - br;
            o[i] = (uint32_t)t;
            br = (uint32_t)(t >> 63);
        }
        return;
    }
    case 21: case 22: {                              // ishl, ishr (logical)
        const uint32_t s = b[0] & (uint32_t)(Fm::W - 1);
        copy<NL>(o, a);
        if (op == 21) {
            shl<NL>(o, s);
        } else {
            uint32_t jam = 0;
            shr_jam<NL>(o, s, jam);
        }
        return;
    }
    case 23:                                         // icmplt, unsigned
        if (ucmp<NL>(a, b) < 0)
            enc_one<NL>(o);
        else
            enc_zero<NL>(o, 0);
        return;
    case 30:                                         // imul: low 32 of low 32s
#pragma unroll
        for (int i = 0; i < NL; i++)
            o[i] = 0;
        o[0] = a[0] * b[0];
        return;
    case 26: {                                       // recip_seed, quiet
        const int sign = a[NL - 1] >> 31;
        const uint32_t ef = (a[NL - 1] >> Fm::TOPF) & Fm::EMASK;
        uint32_t fnz = a[NL - 1] & Fm::FMASK;
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            fnz |= a[i];
        if (ef == Fm::EMASK) {
            if (fnz) enc_qnan<NL>(o); else enc_zero<NL>(o, sign);
            return;
        }
        if (ef == 0) {                               // zero or subnormal
            enc_inf<NL>(o, sign);
            return;
        }
        const int E = (int)ef - Fm::BIAS;
        const int eo = -E - 18;
        if (eo + 17 > Fm::EMAX) {
            enc_inf<NL>(o, sign);
            return;
        }
        seed_pack<NL>(sign, kSeeds.recip[top9<NL>(a)], 18, eo, o);
        return;
    }
    case 27: {                                       // rsqrt_seed, quiet
        const int sign = a[NL - 1] >> 31;
        const uint32_t ef = (a[NL - 1] >> Fm::TOPF) & Fm::EMASK;
        uint32_t fnz = a[NL - 1] & Fm::FMASK;
#pragma unroll
        for (int i = 0; i < NL - 1; i++)
            fnz |= a[i];
        if (ef == Fm::EMASK && fnz) { enc_qnan<NL>(o); return; }
        if (ef == 0) { enc_inf<NL>(o, sign); return; }
        if (sign) { enc_qnan<NL>(o); return; }
        if (ef == Fm::EMASK) { enc_zero<NL>(o, 0); return; }
        const int E = (int)ef - Fm::BIAS;
        const int odd = E & 1;
        const uint32_t r = kSeeds.rsqrt[((uint32_t)odd << 9) | top9<NL>(a)];
        const int nb = 32 - __builtin_clz(r);
        seed_pack<NL>(0, r, nb, -((E - odd) / 2) - 17, o);
        return;
    }
    default:                                         // unassigned
        enc_qnan<NL>(o);
        flags |= FL_INV;
        return;
    }
}

} // namespace fx
