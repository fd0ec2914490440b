// A binary256 FMA spread over eight lanes: lane l of an aligned group of 8
// holds limb l (32 bits, little end first) of each operand and of the
// result, and the operation's wide values (the 474-bit product, the 512-bit
// window of the sum) are held two limbs a lane: limb l and limb l+8.
//
// For a wavefront alone on its SIMD every instruction it issues is time
// (one VALU instruction every ~5 clocks at best), so spreading one
// operation's instructions over eight lanes shortens it: each lane computes
// 8 of the 64 partial products and its two columns of the product, and the
// alignment, the sum and the rounding move limbs between lanes
// (ds_bpermute: ~60 clocks a dependent hop) and resolve carries by ballot
// look-ahead instead of limb-by-limb chains.
//
// The fast path covers the common case and reports when it does not apply:
// a and b normal, c normal or zero, the larger term's leading exponent in
// [EMIN+2, EMAX-2], and a sum whose leading bit stays within one place of
// the larger term's (no deep cancellation).  The result is then normal, and
// the bits and flags are those of fma_core (the same window, the same jam
// and the same rounding: fx_arith.hpp); everything else goes back to the
// caller, which runs fma_core on the whole operands.  Every branch here is
// uniform over a group, and a group's lanes run together.
#pragma once

#include "fx_arith.hpp"

namespace fx {
namespace lp {

using F8 = F<8>;

struct Grp {
    uint32_t l;    // limb index, 0..7
    uint32_t gb;   // the group's first lane in the wavefront
};

__device__ __forceinline__ Grp grp()
{
    const uint32_t lane = __lane_id();
    return Grp{lane & 7u, lane & ~7u};
}

// v as held by lane `src` (0..7) of this lane's group
__device__ __forceinline__ uint32_t from(const Grp &g, uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((g.gb + src) << 2), (int)v);
}

// bit l of the result: p on lane l of this lane's group
__device__ __forceinline__ uint32_t mask8(const Grp &g, bool p)
{
    return (uint32_t)(__builtin_amdgcn_ballot_w64(p) >> g.gb) & 0xFFu;
}

// carries into limbs 0..15 of a sum whose limbs generate (G) or propagate
// (P, disjoint from G) a carry, with cin into limb 0: the carry recurrence
// c[k+1] = G[k] | (P[k] & c[k]) is the one of adding X = G|P and Y = G, so
// the carries are the bits of (X + Y + cin) ^ X ^ Y; bit 16 is the carry out
__device__ __forceinline__ uint32_t carries(uint32_t G, uint32_t P, uint32_t cin)
{
    const uint32_t X = G | P;
    return (X + G + cin) ^ X ^ G;
}

// limb m (0..15, or beyond: zero) of a two-limbs-a-lane value, fetched from
// the lane that holds it: lo/hi are that lane's limbs m&7 and (m&7)+8
__device__ __forceinline__ uint32_t pick(int m, uint32_t lo, uint32_t hi)
{
    return m < 0 ? 0u : m < 8 ? lo : m < 16 ? hi : 0u;
}

// x <<= s (0 <= s < 512), two limbs a lane, zero fill
__device__ __forceinline__ void shl16(const Grp &g, uint32_t &lo, uint32_t &hi, uint32_t s)
{
    const int q = (int)(s >> 5);
    const uint32_t r = s & 31u;
    const int l = (int)g.l;
    const uint32_t s1 = (uint32_t)(l - q) & 7u, s2 = (uint32_t)(l - q - 1) & 7u;
    const uint32_t a_lo = from(g, lo, s1), a_hi = from(g, hi, s1);
    const uint32_t b_lo = from(g, lo, s2), b_hi = from(g, hi, s2);
    // output limb k = l: x[l-q] and x[l-q-1]; k = l+8: x[l+8-q] and x[l+7-q]
    const uint32_t x1 = pick(l - q, a_lo, a_hi), x2 = pick(l - q - 1, b_lo, b_hi);
    const uint32_t y1 = pick(l + 8 - q, a_lo, a_hi), y2 = pick(l + 7 - q, b_lo, b_hi);
    lo = r ? alignbit(x1, x2, 32u - r) : x1;
    hi = r ? alignbit(y1, y2, 32u - r) : y1;
}

// x >>= s (any s), two limbs a lane; returns whether a nonzero bit fell off
__device__ __forceinline__ bool shr16_jam(const Grp &g, uint32_t &lo, uint32_t &hi, uint32_t s)
{
    if (s > 512u)
        s = 512u;
    const int q = (int)(s >> 5);
    const uint32_t r = s & 31u;
    const int l = (int)g.l;
    // the bits that fall off: whole limbs below q, the low r bits of limb q
    const uint32_t part = (1u << r) - 1u;
    const uint32_t off_lo = l < q ? lo : (l == q ? lo & part : 0u);
    const uint32_t off_hi = l + 8 < q ? hi : (l + 8 == q ? hi & part : 0u);
    const bool jam = mask8(g, (off_lo | off_hi) != 0) != 0;
    const uint32_t s1 = (uint32_t)(l + q) & 7u, s2 = (uint32_t)(l + q + 1) & 7u;
    const uint32_t a_lo = from(g, lo, s1), a_hi = from(g, hi, s1);
    const uint32_t b_lo = from(g, lo, s2), b_hi = from(g, hi, s2);
    // output limb k = l: x[l+q], x[l+q+1]; k = l+8: x[l+8+q], x[l+9+q]
    const uint32_t x1 = pick(l + q, a_lo, a_hi), x2 = pick(l + q + 1, b_lo, b_hi);
    const uint32_t y1 = pick(l + 8 + q, a_lo, a_hi), y2 = pick(l + 9 + q, b_lo, b_hi);
    lo = alignbit(x2, x1, r);
    hi = alignbit(y2, y1, r);
    return jam;
}

// x + y + cin, two limbs a lane, carries by look-ahead; returns the carry
// out of limb 15
__device__ __forceinline__ uint32_t add16(const Grp &g, uint32_t &lo, uint32_t &hi,
                                          uint32_t ylo, uint32_t yhi, uint32_t cin)
{
    const uint64_t tl = (uint64_t)lo + ylo, th = (uint64_t)hi + yhi;
    const uint32_t sl = (uint32_t)tl, sh = (uint32_t)th;
    const bool gl = (tl >> 32) != 0, gh = (th >> 32) != 0;
    const uint32_t G = mask8(g, gl) | (mask8(g, gh) << 8);
    const uint32_t P = mask8(g, !gl && sl == 0xFFFFFFFFu) | (mask8(g, !gh && sh == 0xFFFFFFFFu) << 8);
    const uint32_t c = carries(G, P, cin);
    lo = sl + ((c >> g.l) & 1u);
    hi = sh + ((c >> (g.l + 8)) & 1u);
    return (c >> 16) & 1u;
}

// The fast path.  ea, eb, ec are the raw encodings' limbs of this lane; op is
// FMA, ADD, SUB or MUL with ADD/SUB/MUL already steered by the caller to the
// FMA they are defined as (b = 1, c = -c, c = +-0).  On success `out` is
// this lane's limb of the result and `flags` the operation's flags (the
// same on every lane of the group).
__device__ __forceinline__ bool fma8(uint32_t xa, uint32_t xb, uint32_t xc, int rnd,
                                     uint32_t &out, uint32_t &flags)
{
    constexpr int P = F8::P, T = F8::T, BIAS = F8::BIAS, TOPF = F8::TOPF;
    const Grp g = grp();
    const uint32_t a7 = from(g, xa, 7), b7 = from(g, xb, 7), c7 = from(g, xc, 7);
    const int bea = (int)((a7 >> TOPF) & F8::EMASK), beb = (int)((b7 >> TOPF) & F8::EMASK);
    const int bec = (int)((c7 >> TOPF) & F8::EMASK);
    const uint32_t cfrac = g.l == 7 ? (xc & F8::FMASK) : xc;
    const bool czero = bec == 0 && mask8(g, cfrac != 0) == 0;
    const int sa = (int)(a7 >> 31), sb = (int)(b7 >> 31), sc = (int)(c7 >> 31);
    const int sp = sa ^ sb;
    if (bea < 1 || bea > 2 * BIAS || beb < 1 || beb > 2 * BIAS ||
        (!czero && (bec < 1 || bec > 2 * BIAS)))
        return false;
    // significand limbs (the implicit bit in limb 7's bit TOPF)
    const uint32_t top = 1u << TOPF;
    const uint32_t ma = g.l == 7 ? (xa & F8::FMASK) | top : xa;
    const uint32_t mb = g.l == 7 ? (xb & F8::FMASK) | top : xb;
    const uint32_t mc = czero ? 0u : (g.l == 7 ? (xc & F8::FMASK) | top : xc);
    // leading-bit exponents: the product's once its top bit is known
    const int ep = (bea - BIAS - F8::MANW) + (beb - BIAS - F8::MANW);   // lsb of ma*mb
    const int VEc = bec - BIAS;

    // the product: lane l sums columns l and l+8.  With a rotated by m and b
    // limb (8-m) mod 8 broadcast, lane l's product m is a[(l+m)&7]*b[(8-m)&7],
    // of column l (m == 0 or l+m >= 8) or l+8 (otherwise)
    uint64_t acc_lo = 0, acc_hi = 0;
    uint32_t cnt_lo = 0, cnt_hi = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const uint32_t ar = m == 0 ? ma : from(g, ma, (g.l + (uint32_t)m) & 7u);
        const uint32_t bm = from(g, mb, (8u - (uint32_t)m) & 7u);
        const bool tolo = m == 0 || g.l + (uint32_t)m >= 8u;
        mac(acc_lo, cnt_lo, tolo ? ar : 0u, bm);
        mac(acc_hi, cnt_hi, tolo ? 0u : ar, bm);
    }
    // columns to limbs: each column passes its high part to the next, twice
    // (67-bit columns -> 33-bit), then a look-ahead settles the last carries
    uint32_t plo, phi;
    {
        const uint32_t lp = (g.l + 7u) & 7u;           // the lane one down
        const uint32_t l0 = (uint32_t)acc_lo, l1 = (uint32_t)(acc_lo >> 32);
        const uint32_t h0 = (uint32_t)acc_hi, h1 = (uint32_t)(acc_hi >> 32);
        const uint32_t pl1 = from(g, l1, lp), pl2 = from(g, cnt_lo, lp);
        const uint32_t ph1 = from(g, h1, lp), ph2 = from(g, cnt_hi, lp);
        // column l's predecessor is column l-1 (none for l = 0); column l+8's is
        // l+7: lane l-1's high column, or lane 7's low one for l = 0
        const uint32_t ci_lo0 = g.l ? pl1 : 0u, ci_lo1 = g.l ? pl2 : 0u;
        const uint32_t ci_hi0 = g.l ? ph1 : pl1, ci_hi1 = g.l ? ph2 : pl2;
        const uint64_t tl = (uint64_t)l0 + ci_lo0, th = (uint64_t)h0 + ci_hi0;
        const uint32_t kl = (uint32_t)(tl >> 32) + ci_lo1, kh = (uint32_t)(th >> 32) + ci_hi1;
        const uint32_t ul = (uint32_t)tl, uh = (uint32_t)th;
        const uint32_t qkl = from(g, kl, lp), qkh = from(g, kh, lp);
        const uint64_t vl = (uint64_t)ul + (g.l ? qkl : 0u);
        const uint64_t vh = (uint64_t)uh + (g.l ? qkh : qkl);
        plo = (uint32_t)vl;
        phi = (uint32_t)vh;
        const bool gl = (vl >> 32) != 0, gh = (vh >> 32) != 0;
        const uint32_t G = mask8(g, gl) | (mask8(g, gh) << 8);
        const uint32_t Pp = mask8(g, !gl && plo == 0xFFFFFFFFu) |
                            (mask8(g, !gh && phi == 0xFFFFFFFFu) << 8);
        const uint32_t c = carries(G, Pp, 0u);
        plo += (c >> g.l) & 1u;
        phi += (c >> (g.l + 8)) & 1u;
    }
    constexpr int TB = 2 * P - 1;                      // 473: the product's top bit
    const bool ptop = (from(g, phi, (TB / 32) - 8) >> (TB % 32)) & 1u;
    const int VEp = ep + (ptop ? TB : TB - 1);
    // the larger term's leading exponent decides the window and the result's range
    const int d = czero ? -(1 << 20) : VEc - VEp;
    const bool cbig = d > 0;
    const int VEb = cbig ? VEc : VEp;
    if (VEb < F8::EMIN + 2 || VEb > F8::EMAX - 2)
        return false;
    // both terms with their leading bit at T (fma_core's window)
    uint32_t Pl = plo, Ph = phi;
    shl16(g, Pl, Ph, ptop ? (uint32_t)(T - TB) : (uint32_t)(T - TB + 1));
    uint32_t Cl = mc, Ch = 0;
    shl16(g, Cl, Ch, (uint32_t)(T - (P - 1)));
    uint32_t Bl = cbig ? Cl : Pl, Bh = cbig ? Ch : Ph;
    uint32_t Sl = cbig ? Pl : Cl, Sh = cbig ? Ph : Ch;
    const int ad = d < 0 ? -d : d;
    const bool jam = shr16_jam(g, Sl, Sh, (uint32_t)(ad > 512 ? 512 : ad));
    if (g.l == 0 && jam)
        Sl |= 1u;
    const int bsign = cbig ? sc : sp;
    const bool subtract = sc != sp;
    if (subtract) {
        Sl = ~Sl;
        Sh = ~Sh;
    }
    const uint32_t cout = add16(g, Bl, Bh, Sl, Sh, subtract ? 1u : 0u);
    if (subtract && !cout)
        return false;                                  // the smaller term won: rare
    // leading bit at T+1, T or T-1 (limb 15 is lane 7's high limb)
    const uint32_t s15 = from(g, Bh, 7);
    const bool t1 = (s15 >> ((T + 1) % 32)) & 1u, t0 = (s15 >> (T % 32)) & 1u;
    const bool tm = (s15 >> ((T - 1) % 32)) & 1u;
    if (!t1 && !t0 && !tm)
        return false;                                  // deep cancellation
    const int M = t1 ? T + 1 : (t0 ? T : T - 1);
    // K = the P bits from M down (limbs 8.. of the window, one funnel shift)
    constexpr int LI = (T - P) / 32;                   // 8
    static_assert(LI == 8, "the window's rounding limbs are the high ones");
    const uint32_t off = (uint32_t)(M - P + 1 - 32 * LI);   // 11..13
    const uint32_t nxt = from(g, Bh, (g.l + 1u) & 7u);
    uint32_t K = alignbit(g.l == 7 ? 0u : nxt, Bh, off);
    if (g.l == 7)
        K &= (2u << TOPF) - 1u;
    const uint32_t s8 = from(g, Bh, 0);
    const bool gbit = (s8 >> (off - 1)) & 1u;
    const bool sticky = (s8 & ((1u << (off - 1)) - 1u)) != 0 || mask8(g, Bl != 0) != 0;
    int e_norm = VEb + (M - T);
    const bool lsb = from(g, K, 0) & 1u;
    const uint32_t inc = round_incr(rnd, bsign, gbit, sticky, lsb) ? 1u : 0u;
    {
        const uint32_t Pk = mask8(g, K == 0xFFFFFFFFu);
        const uint32_t c = (Pk + inc) ^ Pk;             // carries into limbs 0..7
        K += (c >> g.l) & 1u;
        const bool over = mask8(g, g.l == 7 && (K >> (TOPF + 1)) != 0) != 0;
        if (over) {                                     // 2^P: renormalise, exact
            K = g.l == 7 ? top : 0u;
            e_norm++;
        }
    }
    out = g.l == 7 ? ((uint32_t)bsign << 31) | ((uint32_t)(e_norm + BIAS) << TOPF) |
                         (K & F8::FMASK)
                   : K;
    flags = (gbit || sticky) ? FL_NX : 0u;
    return true;
}

} // namespace lp
} // namespace fx
