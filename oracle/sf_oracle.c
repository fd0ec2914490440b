/* The oracle's arithmetic: a CPU restatement of the reference contract.
 *
 * TEST INFRASTRUCTURE ONLY - see sf_oracle.h.  Every function cites the
 * reference file:line whose behaviour it restates.  The code is our own:
 * exact integer arithmetic on a small fixed bignum, the alignment bounded
 * by substituting a far-smaller term with a representative that lies in
 * the same open rounding interval (argued at oc_fma_finite below).
 */
#include "sf_oracle.h"

#include <pthread.h>
#include <string.h>

/* host/src/softfloat.c:10-16 - the four interchange formats. */
const oc_fmt oc_formats[4] = {
    /* exp_w man_w width prec  bias     emax     emin */
    {      8,   23,   32,  24,    127,     127,    -126 },
    {     11,   52,   64,  53,   1023,    1023,   -1022 },
    {     15,  112,  128, 113,  16383,   16383,  -16382 },
    {     19,  236,  256, 237, 262143,  262143, -262142 },
};

enum { RNE = 0, RTZ = 1, RDN = 2, RUP = 3, RMM = 4 };
enum { K_ZERO, K_SUB, K_NORM, K_INF, K_NAN };

/* ---- bignum ------------------------------------------------------------ */

static void bz(oc_big *r) { r->n = 0; }

static void bnorm(oc_big *r)
{
    while (r->n > 0 && r->w[r->n - 1] == 0)
        r->n--;
}

static void bset32(oc_big *r, uint32_t x)
{
    r->w[0] = x;
    r->n = x ? 1 : 0;
}

static void bcopy_(oc_big *r, const oc_big *a)
{
    if (r != a)
        memcpy(r, a, sizeof *r);
}

static int bbitlen(const oc_big *a)
{
    uint32_t t;
    int b = 0;
    if (a->n == 0)
        return 0;
    t = a->w[a->n - 1];
    while (t) { b++; t >>= 1; }
    return (a->n - 1) * 32 + b;
}

static int bbit(const oc_big *a, int i)
{
    if (i < 0 || (i >> 5) >= a->n)
        return 0;
    return (int)((a->w[i >> 5] >> (i & 31)) & 1u);
}

static void bsetbit(oc_big *r, int i)
{
    int l = i >> 5;
    while (r->n <= l)
        r->w[r->n++] = 0;
    r->w[l] |= 1u << (i & 31);
}

static void bclearbit(oc_big *r, int i)
{
    if ((i >> 5) < r->n) {
        r->w[i >> 5] &= ~(1u << (i & 31));
        bnorm(r);
    }
}

/* keep bits [0, nbits) */
static void bmask(oc_big *r, int nbits)
{
    int l = nbits >> 5, b = nbits & 31;
    if (nbits <= 0) { r->n = 0; return; }
    if (l >= r->n)
        return;
    if (b) {
        r->w[l] &= (1u << b) - 1u;
        r->n = l + 1;
    } else {
        r->n = l;
    }
    bnorm(r);
}

/* any of bits [0, k) set */
static int blow_nonzero(const oc_big *a, int k)
{
    int i;
    for (i = 0; i < a->n && i * 32 < k; i++) {
        uint32_t m = (k - i * 32 >= 32) ? 0xFFFFFFFFu
                                         : ((1u << (k - i * 32)) - 1u);
        if (a->w[i] & m)
            return 1;
    }
    return 0;
}

static void bshl(oc_big *r, const oc_big *a, int k)
{
    oc_big t;
    int ls = k >> 5, bs = k & 31, i;
    if (a->n == 0) { bz(r); return; }
    memset(&t, 0, sizeof t);
    for (i = a->n - 1; i >= 0; i--) {
        uint64_t v = (uint64_t)a->w[i] << bs;
        if (i + ls >= OC_LIMBS)
            continue;                   /* past the container: dropped */
        if (i + ls + 1 < OC_LIMBS)
            t.w[i + ls + 1] |= (uint32_t)(v >> 32);
        t.w[i + ls] |= (uint32_t)v;
    }
    t.n = a->n + ls + 1;
    if (t.n > OC_LIMBS)
        t.n = OC_LIMBS;
    bnorm(&t);
    bcopy_(r, &t);
}

static void bshr(oc_big *r, const oc_big *a, int k)
{
    oc_big t;
    int ls = k >> 5, bs = k & 31, i;
    memset(&t, 0, sizeof t);
    for (i = ls; i < a->n; i++) {
        uint64_t v = a->w[i];
        if (i + 1 < a->n)
            v |= (uint64_t)a->w[i + 1] << 32;
        t.w[i - ls] = (uint32_t)(v >> bs);
    }
    t.n = a->n - ls > 0 ? a->n - ls : 0;
    bnorm(&t);
    bcopy_(r, &t);
}

static int bcmp_(const oc_big *a, const oc_big *b)
{
    int i;
    if (a->n != b->n)
        return a->n < b->n ? -1 : 1;
    for (i = a->n - 1; i >= 0; i--)
        if (a->w[i] != b->w[i])
            return a->w[i] < b->w[i] ? -1 : 1;
    return 0;
}

static void badd(oc_big *r, const oc_big *a, const oc_big *b)
{
    oc_big t;
    uint64_t c = 0;
    int n = a->n > b->n ? a->n : b->n, i;
    for (i = 0; i < n; i++) {
        c += (uint64_t)(i < a->n ? a->w[i] : 0) + (i < b->n ? b->w[i] : 0);
        t.w[i] = (uint32_t)c;
        c >>= 32;
    }
    if (c && n < OC_LIMBS)
        t.w[n++] = (uint32_t)c;
    t.n = n;
    bnorm(&t);
    bcopy_(r, &t);
}

/* r = a - b, a >= b */
static void bsub(oc_big *r, const oc_big *a, const oc_big *b)
{
    oc_big t;
    int64_t c = 0;
    int i;
    for (i = 0; i < a->n; i++) {
        c += (int64_t)a->w[i] - (int64_t)(i < b->n ? b->w[i] : 0);
        t.w[i] = (uint32_t)c;
        c = c < 0 ? -1 : 0;
        /* arithmetic on the borrow, not a shift of a negative value */
    }
    t.n = a->n;
    bnorm(&t);
    bcopy_(r, &t);
}

static void bmul(oc_big *r, const oc_big *a, const oc_big *b)
{
    oc_big t;
    int i, j;
    memset(&t, 0, sizeof t);
    for (i = 0; i < a->n; i++) {
        uint64_t c = 0;
        for (j = 0; j < b->n && i + j < OC_LIMBS; j++) {
            c += (uint64_t)a->w[i] * b->w[j] + t.w[i + j];
            t.w[i + j] = (uint32_t)c;
            c >>= 32;
        }
        if (i + j < OC_LIMBS)
            t.w[i + j] = (uint32_t)c;
    }
    t.n = a->n + b->n < OC_LIMBS ? a->n + b->n : OC_LIMBS;
    bnorm(&t);
    bcopy_(r, &t);
}

/* q = floor(a / b), rem_nz = (a mod b != 0); b != 0.  Bit-serial: the
 * oracle wants obviously-right, not fast. */
static void bdivmod(oc_big *q, int *rem_nz, const oc_big *a, const oc_big *b)
{
    oc_big r, t;
    int i;
    bz(&r);
    bz(&t);
    for (i = bbitlen(a) - 1; i >= 0; i--) {
        bshl(&r, &r, 1);
        if (bbit(a, i))
            bsetbit(&r, 0);
        if (bcmp_(&r, b) >= 0) {
            bsub(&r, &r, b);
            bsetbit(&t, i);
        }
    }
    bcopy_(q, &t);
    *rem_nz = r.n != 0;
}

/* s = floor(sqrt(a)), exact_ = (s*s == a).  Digit-by-digit, base 2. */
static void bisqrt(oc_big *s, int *exact_, const oc_big *a)
{
    oc_big rem, root, trial;
    int i, top = (bbitlen(a) + 1) / 2;
    bz(&rem);
    bz(&root);
    for (i = top - 1; i >= 0; i--) {
        /* bring down two bits */
        bshl(&rem, &rem, 2);
        if (bbit(a, 2 * i + 1)) bsetbit(&rem, 1);
        if (bbit(a, 2 * i))     bsetbit(&rem, 0);
        /* trial = 4*root + 1 */
        bshl(&trial, &root, 2);
        bsetbit(&trial, 0);
        bshl(&root, &root, 1);
        if (bcmp_(&rem, &trial) >= 0) {
            bsub(&rem, &rem, &trial);
            bsetbit(&root, 0);
        }
    }
    bcopy_(s, &root);
    *exact_ = rem.n == 0;
}

void oc_load(oc_big *r, const uint8_t *le, int nbytes)
{
    int i;
    memset(r, 0, sizeof *r);
    for (i = 0; i < nbytes; i++)
        r->w[i >> 2] |= (uint32_t)le[i] << (8 * (i & 3));
    r->n = (nbytes + 3) / 4;
    bnorm(r);
}

void oc_store(const oc_big *a, uint8_t *le, int nbytes)
{
    int i;
    for (i = 0; i < nbytes; i++)
        le[i] = (i >> 2) < a->n ? (uint8_t)(a->w[i >> 2] >> (8 * (i & 3))) : 0;
}

/* ---- encodings ----------------------------------------------------------- */

static void set_field(oc_big *r, int lo, int nbits, uint32_t v)
{
    int i;
    for (i = 0; i < nbits; i++)
        if ((v >> i) & 1u)
            bsetbit(r, lo + i);
}

static uint32_t exp_field(const oc_fmt *f, const oc_big *x)
{
    uint32_t e = 0;
    int i;
    for (i = 0; i < f->exp_w; i++)
        e |= (uint32_t)bbit(x, f->man_w + i) << i;
    return e;
}

static uint32_t exp_all_ones(const oc_fmt *f)
{
    return (1u << f->exp_w) - 1u;
}

static void enc_zero(const oc_fmt *f, int sign, oc_big *o)
{
    bz(o);
    if (sign) bsetbit(o, f->width - 1);
}

static void enc_inf(const oc_fmt *f, int sign, oc_big *o)
{
    bz(o);
    set_field(o, f->man_w, f->exp_w, exp_all_ones(f));
    if (sign) bsetbit(o, f->width - 1);
}

/* host/src/softfloat.c:50-56: the only NaN arithmetic ever produces. */
static void enc_qnan(const oc_fmt *f, oc_big *o)
{
    bz(o);
    set_field(o, f->man_w, f->exp_w, exp_all_ones(f));
    bsetbit(o, f->man_w - 1);
}

static void enc_maxnorm(const oc_fmt *f, int sign, oc_big *o)
{
    int i;
    bz(o);
    for (i = 0; i < f->man_w; i++)
        bsetbit(o, i);
    set_field(o, f->man_w, f->exp_w, exp_all_ones(f) - 1u);
    if (sign) bsetbit(o, f->width - 1);
}

static void enc_one(const oc_fmt *f, oc_big *o)
{
    bz(o);
    set_field(o, f->man_w, f->exp_w, (uint32_t)f->bias);
}

typedef struct {
    int kind, sign, snan, e;
    oc_big m;       /* value = (-1)^sign * m * 2^e for finite kinds */
} unp;

/* host/src/softfloat.c:85-116 */
static void unpack(const oc_fmt *f, const oc_big *x, unp *u)
{
    uint32_t ef = exp_field(f, x);
    oc_big frac;
    bcopy_(&frac, x);
    bmask(&frac, f->man_w);
    u->sign = bbit(x, f->width - 1);
    u->snan = 0;
    u->e = 0;
    bz(&u->m);
    if (ef == exp_all_ones(f)) {
        if (frac.n == 0) {
            u->kind = K_INF;
        } else {
            u->kind = K_NAN;
            u->snan = !bbit(x, f->man_w - 1);
        }
        return;
    }
    if (ef == 0) {
        if (frac.n == 0) {
            u->kind = K_ZERO;
            return;
        }
        u->kind = K_SUB;
        bcopy_(&u->m, &frac);
        u->e = f->emin - f->man_w;
        return;
    }
    u->kind = K_NORM;
    bcopy_(&u->m, &frac);
    bsetbit(&u->m, f->man_w);
    u->e = (int)ef - f->bias - f->man_w;
}

/* ---- rounding ------------------------------------------------------------ */

/* Negative controls (tests/test_negative_controls.py).  A build with
 * -DORACLE_PLANT=k breaks one rule on purpose so that the comparison with
 * the reference-written golden vectors must fail: 1 RMM rounds as RNE,
 * 2 UNDERFLOW is dropped for tiny inexact results, 3 RNE ties go to odd,
 * 4 tininess is decided before rounding.  The default build (and every
 * build the checks use) has none of them. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

/* host/src/softfloat.c:120-134: does a mode increment, given the sign,
 * the first discarded bit, whether anything below it is set, and the
 * kept lsb. */
static int round_incr(int rnd, int sign, int guard, int sticky, int lsb)
{
    switch (rnd) {
#if ORACLE_PLANT == 3
    case RNE: return guard && (sticky || !lsb);
#else
    case RNE: return guard && (sticky || lsb);
#endif
    case RTZ: return 0;
    case RDN: return sign && (guard || sticky);
    case RUP: return !sign && (guard || sticky);
#if ORACLE_PLANT == 1
    default:  return guard && (sticky || lsb);
#else
    default:  return guard;                       /* RMM */
#endif
    }
}

/* host/src/softfloat.c:141-149: overflow delivers inf or the largest
 * finite, by mode and sign. */
static int overflow_to_inf(int rnd, int sign)
{
    if (rnd == RTZ) return 0;
    if (rnd == RDN) return sign;
    if (rnd == RUP) return !sign;
    return 1;
}

/* m * 2^e rounded to a multiple of 2^q: kept = the multiple / 2^q. */
static void round_to(const oc_big *m, int e, int q, int sign, int rnd,
                     oc_big *kept, int *inexact)
{
    int sh = q - e;
    *inexact = 0;
    if (sh <= 0) {
        bshl(kept, m, -sh);
        return;
    }
    bshr(kept, m, sh);
    if (!blow_nonzero(m, sh))
        return;
    *inexact = 1;
    if (round_incr(rnd, sign, bbit(m, sh - 1), blow_nonzero(m, sh - 1),
                   bbit(kept, 0))) {
        oc_big one;
        bset32(&one, 1);
        badd(kept, kept, &one);
    }
}

/* host/src/softfloat.c:165-261: the exact non-zero value
 * (-1)^sign * m * 2^e into the format.  Tininess is decided after
 * rounding: the value rounded to p bits with an unbounded exponent range
 * is below 2^emin.  UNDERFLOW is raised for a tiny inexact result and for
 * a result that rounds to zero; OVERFLOW always with INEXACT. */
static void round_pack(const oc_fmt *f, int sign, const oc_big *m, int e,
                       int rnd, oc_big *out, uint32_t *flags)
{
    int p = f->prec, L = bbitlen(m), e_norm, q, inexact, dummy, e_res, tiny;
    oc_big ku, k;
    uint32_t fl = 0;

    e_norm = e + L - 1;
    round_to(m, e, e_norm - (p - 1), sign, rnd, &ku, &dummy);
#if ORACLE_PLANT == 4
    tiny = e_norm < f->emin;
#else
    tiny = (e_norm - (p - 1) + bbitlen(&ku) - 1) < f->emin;
#endif

    q = (e_norm > f->emin ? e_norm : f->emin) - (p - 1);
    round_to(m, e, q, sign, rnd, &k, &inexact);
    if (inexact)
        fl |= OC_FLAG_INEXACT;
    if (k.n == 0) {
        enc_zero(f, sign, out);
        *flags = fl | OC_FLAG_UNDERFLOW;
        return;
    }
    e_res = q + bbitlen(&k) - 1;
    if (e_res > f->emax) {
        if (overflow_to_inf(rnd, sign))
            enc_inf(f, sign, out);
        else
            enc_maxnorm(f, sign, out);
        *flags = fl | OC_FLAG_OVERFLOW | OC_FLAG_INEXACT;
        return;
    }
    if (tiny && inexact && ORACLE_PLANT != 2)
        fl |= OC_FLAG_UNDERFLOW;
    if (e_res < f->emin) {
        /* subnormal: k's lsb weighs 2^(emin - p + 1), so k is the field */
        bcopy_(out, &k);
    } else {
        if (bbitlen(&k) == p + 1)
            bshr(&k, &k, 1);                  /* carried to 2^p: exact */
        bclearbit(&k, f->man_w);
        bcopy_(out, &k);
        set_field(out, f->man_w, f->exp_w, (uint32_t)(e_res + f->bias));
    }
    if (sign)
        bsetbit(out, f->width - 1);
    *flags = fl;
}

/* ---- fused multiply-add -------------------------------------------------- */

static int is_nan_enc(const oc_fmt *f, const oc_big *x, int *snan)
{
    oc_big frac;
    if (exp_field(f, x) != exp_all_ones(f))
        return 0;
    bcopy_(&frac, x);
    bmask(&frac, f->man_w);
    if (frac.n == 0)
        return 0;
    if (snan)
        *snan = !bbit(x, f->man_w - 1);
    return 1;
}

/* The finite, non-zero a*b + c: exact sum, one rounding.
 *
 * Alignment bound.  P = mp*2^ep (at most 2p bits), C = mc*2^ec (at most p
 * bits), VE the exponent of each term's leading bit.  If the smaller term
 * leads more than 2p+4 binades below the larger, it is below 2^(VEbig-2p-3).
 * Every rounding boundary of the result (representable values and midpoints
 * at the rounding position, with or without the exponent bound) and the
 * larger term itself are multiples of 2^g with g >= VEbig - 2p + 1, so the
 * open interval between the larger term and its neighbour on that grid
 * holds no boundary: any tiny term of the same sign in (0, 2^g) gives the
 * same rounded result and the same flags.  We substitute 2^(VEbig-2p-3).
 * Otherwise the exact sum spans at most 5p + 4 bits. */
static void fma_finite(const oc_fmt *f, int sp, const oc_big *mp, int ep,
                       int sc, const oc_big *mc, int ec, int rnd,
                       oc_big *out, uint32_t *flags)
{
    const int T = 2 * f->prec + 4;
    int VEp = ep + bbitlen(mp) - 1, VEc = ec + bbitlen(mc) - 1, e0, c;
    oc_big tp, tc, s, one;

    bset32(&one, 1);
    if (VEp - VEc > T) {
        mc = &one;
        ec = VEp - 2 * f->prec - 3;
    } else if (VEc - VEp > T) {
        mp = &one;
        ep = VEc - 2 * f->prec - 3;
    }
    e0 = ep < ec ? ep : ec;
    bshl(&tp, mp, ep - e0);
    bshl(&tc, mc, ec - e0);
    if (sp == sc) {
        badd(&s, &tp, &tc);
        round_pack(f, sp, &s, e0, rnd, out, flags);
        return;
    }
    c = bcmp_(&tp, &tc);
    if (c == 0) {
        /* host/src/softfloat.c:153-156, 416-419: exact cancellation is +0
         * in every mode but roundTowardNegative */
        enc_zero(f, rnd == RDN, out);
        *flags = 0;
        return;
    }
    if (c > 0) {
        bsub(&s, &tp, &tc);
        round_pack(f, sp, &s, e0, rnd, out, flags);
    } else {
        bsub(&s, &tc, &tp);
        round_pack(f, sc, &s, e0, rnd, out, flags);
    }
}

/* host/src/softfloat.c:324-428 - the special-case order is the
 * contract's: NaN, then inf*0, inf-inf, inf, c inf, a zero factor, and
 * only then the product. */
static void fma_(const oc_fmt *f, const oc_big *xa, const oc_big *xb,
                 const oc_big *xc, int rnd, oc_big *out, uint32_t *flags)
{
    unp ua, ub, uc;
    int sp;
    oc_big mp;

    unpack(f, xa, &ua);
    unpack(f, xb, &ub);
    unpack(f, xc, &uc);
    *flags = 0;
    if (ua.kind == K_NAN || ub.kind == K_NAN || uc.kind == K_NAN) {
        enc_qnan(f, out);
        *flags = (ua.snan || ub.snan || uc.snan) ? OC_FLAG_INVALID : 0;
        return;
    }
    sp = ua.sign ^ ub.sign;
    if (ua.kind == K_INF || ub.kind == K_INF) {
        if (ua.kind == K_ZERO || ub.kind == K_ZERO ||
            (uc.kind == K_INF && uc.sign != sp)) {
            enc_qnan(f, out);
            *flags = OC_FLAG_INVALID;
            return;
        }
        enc_inf(f, sp, out);
        return;
    }
    if (uc.kind == K_INF) {
        enc_inf(f, uc.sign, out);
        return;
    }
    if (ua.kind == K_ZERO || ub.kind == K_ZERO) {
        if (uc.kind == K_ZERO)
            enc_zero(f, uc.sign == sp ? uc.sign : (rnd == RDN), out);
        else
            bcopy_(out, xc);            /* 0*b + c is c, bits and all */
        return;
    }
    bmul(&mp, &ua.m, &ub.m);
    if (uc.kind == K_ZERO) {
        round_pack(f, sp, &mp, ua.e + ub.e, rnd, out, flags);
        return;
    }
    fma_finite(f, sp, &mp, ua.e + ub.e, uc.sign, &uc.m, uc.e, rnd, out,
               flags);
}

/* ---- the other operations ------------------------------------------------ */

static void magnitude(const oc_fmt *f, const oc_big *x, oc_big *r)
{
    bcopy_(r, x);
    bclearbit(r, f->width - 1);
}

/* x < y, neither a NaN, -0 == +0 (host/src/softfloat.c:487-503) */
static int numeric_lt(const oc_fmt *f, const oc_big *x, const oc_big *y)
{
    oc_big mx, my;
    int sx = bbit(x, f->width - 1), sy = bbit(y, f->width - 1);
    magnitude(f, x, &mx);
    magnitude(f, y, &my);
    if (mx.n == 0 && my.n == 0)
        return 0;
    if (sx != sy)
        return sx;
    return sx ? bcmp_(&mx, &my) > 0 : bcmp_(&mx, &my) < 0;
}

/* host/src/softfloat.c:521-564: 754-2019 9.6 minimum/maximum and the
 * ...Number forms. */
static void minmax(const oc_fmt *f, const oc_big *a, const oc_big *b,
                   int want_max, int number, oc_big *out, uint32_t *flags)
{
    int sa = 0, sb = 0;
    int an = is_nan_enc(f, a, &sa), bn = is_nan_enc(f, b, &sb);
    oc_big ma, mb;
    *flags = ((an && sa) || (bn && sb)) ? OC_FLAG_INVALID : 0;
    if (an || bn) {
        if (!number || (an && bn))
            enc_qnan(f, out);
        else
            bcopy_(out, an ? b : a);
        return;
    }
    magnitude(f, a, &ma);
    magnitude(f, b, &mb);
    if (ma.n == 0 && mb.n == 0) {
        int za = bbit(a, f->width - 1), zb = bbit(b, f->width - 1);
        if (za == zb)
            bcopy_(out, a);
        else
            bcopy_(out, (want_max ? !za : za) ? a : b);
        return;
    }
    if (want_max)
        bcopy_(out, numeric_lt(f, a, b) ? b : a);
    else
        bcopy_(out, numeric_lt(f, a, b) ? a : b);
}

/* host/src/softfloat.c:569-591: quiet comparison to 1.0 or +0. */
static void compare(const oc_fmt *f, const oc_big *a, const oc_big *b,
                    int op, oc_big *out, uint32_t *flags)
{
    int sa = 0, sb = 0, lt, gt, t;
    int an = is_nan_enc(f, a, &sa), bn = is_nan_enc(f, b, &sb);
    *flags = ((an && sa) || (bn && sb)) ? OC_FLAG_INVALID : 0;
    if (an || bn) {
        enc_zero(f, 0, out);
        return;
    }
    lt = numeric_lt(f, a, b);
    gt = numeric_lt(f, b, a);
    t = op == 12 ? lt : op == 13 ? (lt || !gt) : (!lt && !gt);
    if (t)
        enc_one(f, out);
    else
        enc_zero(f, 0, out);
}

/* host/src/softfloat.c:597-647: the integer group on width-bit words. */
static void integer_op(const oc_fmt *f, int op, const oc_big *a,
                       const oc_big *b, oc_big *out)
{
    int i, sh = (b->n ? (int)b->w[0] : 0) & (f->width - 1);
    oc_big t;
    switch (op) {
    case 16: case 17: case 18:
        memset(&t, 0, sizeof t);
        for (i = 0; i < OC_LIMBS; i++) {
            uint32_t x = i < a->n ? a->w[i] : 0, y = i < b->n ? b->w[i] : 0;
            t.w[i] = op == 16 ? (x & y) : op == 17 ? (x | y) : (x ^ y);
        }
        t.n = OC_LIMBS;
        bnorm(&t);
        bcopy_(out, &t);
        return;
    case 19:
        badd(out, a, b);
        bmask(out, f->width);
        return;
    case 20:
        bcopy_(&t, a);
        bsetbit(&t, f->width);
        bsub(out, &t, b);
        bmask(out, f->width);
        return;
    case 21:
        bshl(out, a, sh);
        bmask(out, f->width);
        return;
    case 22:
        bshr(out, a, sh);
        return;
    case 23:
        if (bcmp_(a, b) < 0) enc_one(f, out); else enc_zero(f, 0, out);
        return;
    default: {                                  /* 30, imul */
        uint32_t x = a->n ? a->w[0] : 0, y = b->n ? b->w[0] : 0;
        bset32(out, x * y);
        return;
    }
    }
}

/* The reciprocal seed table entry: the nearest integer to
 * 2^18 / (1 + (i + 1/2) / 2^9) = 2^28 / (2^10 + 2i + 1), ties to even
 * (host/src/softfloat.c:663-680 defines the same value). */
static uint32_t recip_entry(uint32_t i)
{
    uint64_t den = 1024u + 2u * i + 1u, num = 1ull << 28;
    uint64_t q = num / den, r = num % den;
    if (2 * r > den || (2 * r == den && (q & 1)))
        q++;
    return (uint32_t)q;
}

/* The reciprocal-sqrt seed table entry: the nearest integer to
 * 2^17 / sqrt(mid), mid = ((2^10 + 2i + 1) / 2^10) * 2^odd, i.e. to
 * sqrt(2^44 / m2) with m2 = (2^10 + 2i + 1) << odd; rounded up exactly
 * when (q + 1/2)^2 * m2 < 2^44 (host/src/softfloat.c:684-705). */
static uint32_t rsqrt_entry(uint32_t j)
{
    uint32_t odd = j >> 9, i = j & 511u;
    uint64_t m2 = (uint64_t)(1024u + 2u * i + 1u) << odd;
    uint64_t lo = 0, hi = 1u << 18;
    while (hi - lo > 1) {                   /* largest q, q^2 m2 <= 2^44 */
        uint64_t mid = (lo + hi) / 2;
        if (mid * mid * m2 <= (1ull << 44)) lo = mid; else hi = mid;
    }
    if ((2 * lo + 1) * (2 * lo + 1) * m2 < (1ull << 46))
        lo++;
    return (uint32_t)lo;
}

/* host/src/softfloat.c:717-748: the quiet reciprocal seed. */
static void recip_seed(const oc_fmt *f, const oc_big *x, oc_big *out)
{
    int sign = bbit(x, f->width - 1);
    uint32_t ef = exp_field(f, x), dummy;
    oc_big frac, m;
    int E, eo;
    bcopy_(&frac, x);
    bmask(&frac, f->man_w);
    if (ef == exp_all_ones(f)) {
        if (frac.n) enc_qnan(f, out); else enc_zero(f, sign, out);
        return;
    }
    if (ef == 0) {                          /* zero or subnormal */
        enc_inf(f, sign, out);
        return;
    }
    E = (int)ef - f->bias;
    {
        oc_big top;
        bshr(&top, x, f->man_w - 9);
        bmask(&top, 9);
        bset32(&m, recip_entry(top.n ? top.w[0] : 0));
    }
    eo = -E - 18;
    if (eo + 17 > f->emax) {
        enc_inf(f, sign, out);
        return;
    }
    round_pack(f, sign, &m, eo, RNE, out, &dummy);
}

/* host/src/softfloat.c:750-787: the quiet reciprocal-sqrt seed. */
static void rsqrt_seed(const oc_fmt *f, const oc_big *x, oc_big *out)
{
    int sign = bbit(x, f->width - 1);
    uint32_t ef = exp_field(f, x), dummy;
    oc_big frac, m, top;
    int E, odd, eo;
    bcopy_(&frac, x);
    bmask(&frac, f->man_w);
    if (ef == exp_all_ones(f) && frac.n) { enc_qnan(f, out); return; }
    if (ef == 0) { enc_inf(f, sign, out); return; }
    if (sign) { enc_qnan(f, out); return; }
    if (ef == exp_all_ones(f)) { enc_zero(f, 0, out); return; }
    E = (int)ef - f->bias;
    odd = E & 1;
    bshr(&top, x, f->man_w - 9);
    bmask(&top, 9);
    bset32(&m, rsqrt_entry(((uint32_t)odd << 9) | (top.n ? top.w[0] : 0)));
    eo = -((E - odd) / 2) - 17;
    round_pack(f, 0, &m, eo, RNE, out, &dummy);
}

/* host/src/softfloat.c:939-1009 with the steering of 438-468. */
void oc_compute_big(const oc_fmt *f, int op, int rnd, const oc_big *a,
                    const oc_big *b, const oc_big *c, oc_big *out,
                    uint32_t *flags)
{
    oc_big sb, sc;
    *flags = 0;
    switch (op) {
    case 0:
        fma_(f, a, b, c, rnd, out, flags);
        return;
    case 1:                                     /* a*1 + c */
    case 2:                                     /* a*1 + (-c) */
        enc_one(f, &sb);
        bcopy_(&sc, c);
        if (op == 2) {
            if (bbit(&sc, f->width - 1)) bclearbit(&sc, f->width - 1);
            else bsetbit(&sc, f->width - 1);
        }
        fma_(f, a, &sb, &sc, rnd, out, flags);
        return;
    case 3:                                     /* a*b + (+-0) */
        enc_zero(f, bbit(a, f->width - 1) ^ bbit(b, f->width - 1), &sc);
        fma_(f, a, b, &sc, rnd, out, flags);
        return;
    case 4:
        magnitude(f, a, out);
        return;
    case 5:
        bcopy_(out, a);
        if (bbit(a, f->width - 1)) bclearbit(out, f->width - 1);
        else bsetbit(out, f->width - 1);
        return;
    case 6:
        magnitude(f, a, out);
        if (bbit(b, f->width - 1)) bsetbit(out, f->width - 1);
        return;
    case 7:  minmax(f, a, b, 0, 0, out, flags); return;
    case 8:  minmax(f, a, b, 1, 0, out, flags); return;
    case 9:  minmax(f, a, b, 0, 1, out, flags); return;
    case 10: minmax(f, a, b, 1, 1, out, flags); return;
    case 11: {
        oc_big mc;
        magnitude(f, c, &mc);
        bcopy_(out, mc.n ? a : b);
        return;
    }
    case 12: case 13: case 14:
        compare(f, a, b, op, out, flags);
        return;
    case 16: case 17: case 18: case 19: case 20: case 21: case 22: case 23:
    case 30:
        integer_op(f, op, a, b, out);
        return;
    case 26: recip_seed(f, a, out); return;
    case 27: rsqrt_seed(f, a, out); return;
    default:
        /* unassigned opcodes (and the reductions, which have no
         * elementwise meaning) answer the canonical qNaN and invalid */
        enc_qnan(f, out);
        *flags = OC_FLAG_INVALID;
        return;
    }
}

/* ---- division and square root -------------------------------------------
 * The correctly-rounded contract of python/cft_golden/softfloat.py:440-506
 * (read as text): long division / integer square root carried to at least
 * p+3 bits, the remainder folded into one appended sticky bit two or more
 * places below the rounding position, then the one round_pack. */

static void div_(const oc_fmt *f, const oc_big *xa, const oc_big *xb, int rnd,
                 oc_big *out, uint32_t *flags)
{
    unp ua, ub;
    int sq, k, rem_nz;
    oc_big num, q;
    unpack(f, xa, &ua);
    unpack(f, xb, &ub);
    *flags = 0;
    if (ua.kind == K_NAN || ub.kind == K_NAN) {
        enc_qnan(f, out);
        *flags = (ua.snan || ub.snan) ? OC_FLAG_INVALID : 0;
        return;
    }
    sq = ua.sign ^ ub.sign;
    if (ua.kind == K_INF) {
        if (ub.kind == K_INF) { enc_qnan(f, out); *flags = OC_FLAG_INVALID; }
        else enc_inf(f, sq, out);
        return;
    }
    if (ub.kind == K_INF) { enc_zero(f, sq, out); return; }
    if (ub.kind == K_ZERO) {
        if (ua.kind == K_ZERO) { enc_qnan(f, out); *flags = OC_FLAG_INVALID; }
        else { enc_inf(f, sq, out); *flags = OC_FLAG_DIVZERO; }
        return;
    }
    if (ua.kind == K_ZERO) { enc_zero(f, sq, out); return; }
    k = (f->prec + 3) + bbitlen(&ub.m) - bbitlen(&ua.m) + 1;
    if (k < 0) k = 0;
    bshl(&num, &ua.m, k);
    bdivmod(&q, &rem_nz, &num, &ub.m);
    bshl(&q, &q, 1);
    if (rem_nz) bsetbit(&q, 0);
    round_pack(f, sq, &q, ua.e - ub.e - k - 1, rnd, out, flags);
}

static void sqrt_(const oc_fmt *f, const oc_big *xa, int rnd, oc_big *out,
                  uint32_t *flags)
{
    unp ua;
    int t, e, exact_;
    oc_big m, r;
    unpack(f, xa, &ua);
    *flags = 0;
    if (ua.kind == K_NAN) {
        enc_qnan(f, out);
        *flags = ua.snan ? OC_FLAG_INVALID : 0;
        return;
    }
    if (ua.kind == K_ZERO) { bcopy_(out, xa); return; }   /* +-0 itself */
    if (ua.sign) { enc_qnan(f, out); *flags = OC_FLAG_INVALID; return; }
    if (ua.kind == K_INF) { bcopy_(out, xa); return; }
    t = 2 * (f->prec + 3) - bbitlen(&ua.m);
    if (t < 0) t = 0;
    if ((t ^ ua.e) & 1) t++;               /* keep e - t even */
    bshl(&m, &ua.m, t);
    e = ua.e - t;
    bisqrt(&r, &exact_, &m);
    bshl(&r, &r, 1);
    if (!exact_) bsetbit(&r, 0);
    /* e is even, so e / 2 is exact whatever its sign */
    round_pack(f, 0, &r, e / 2 - 1, rnd, out, flags);
}

/* ---- entry points -------------------------------------------------------- */

static int fmt_ok(int fmt) { return fmt >= 0 && fmt <= 3; }

int oc_compute(int fmt, int op, int rnd, const uint8_t *a, const uint8_t *b,
               const uint8_t *c, uint8_t *out, uint32_t *flags)
{
    const oc_fmt *f;
    oc_big ba, bb, bc, bo;
    int esz;
    if (!fmt_ok(fmt)) return 1;
    f = &oc_formats[fmt];
    esz = f->width / 8;
    bz(&ba); bz(&bb); bz(&bc);
    if (a) oc_load(&ba, a, esz);
    if (b) oc_load(&bb, b, esz);
    if (c) oc_load(&bc, c, esz);
    oc_compute_big(f, op, rnd, &ba, &bb, &bc, &bo, flags);
    oc_store(&bo, out, esz);
    return 0;
}

int oc_div(int fmt, int rnd, const uint8_t *a, const uint8_t *b, uint8_t *out,
           uint32_t *flags)
{
    const oc_fmt *f;
    oc_big ba, bb, bo;
    if (!fmt_ok(fmt)) return 1;
    f = &oc_formats[fmt];
    oc_load(&ba, a, f->width / 8);
    oc_load(&bb, b, f->width / 8);
    div_(f, &ba, &bb, rnd, &bo, flags);
    oc_store(&bo, out, f->width / 8);
    return 0;
}

int oc_sqrt(int fmt, int rnd, const uint8_t *a, uint8_t *out, uint32_t *flags)
{
    const oc_fmt *f;
    oc_big ba, bo;
    if (!fmt_ok(fmt)) return 1;
    f = &oc_formats[fmt];
    oc_load(&ba, a, f->width / 8);
    sqrt_(f, &ba, rnd, &bo, flags);
    oc_store(&bo, out, f->width / 8);
    return 0;
}

int oc_round_pack(int fmt, int sign, const uint8_t *m, int mbytes, int e,
                  int rnd, uint8_t *out, uint32_t *flags)
{
    const oc_fmt *f;
    oc_big bm, bo;
    if (!fmt_ok(fmt) || mbytes > OC_LIMBS * 4) return 1;
    f = &oc_formats[fmt];
    oc_load(&bm, m, mbytes);
    if (bm.n == 0) return 1;
    round_pack(f, sign, &bm, e, rnd, &bo, flags);
    oc_store(&bo, out, f->width / 8);
    return 0;
}

typedef struct {
    int fmt, op, rnd;
    const uint8_t *a, *b, *c;
    uint8_t *d, *lane_flags;
    size_t lo, hi;
    uint32_t smask, flags;
} run_job;

static void *run_slice(void *arg)
{
    run_job *J = (run_job *)arg;
    const size_t esz = (size_t)oc_formats[J->fmt].width / 8;
    const size_t sa = (J->smask & 1u) ? 0 : esz, sb = (J->smask & 2u) ? 0 : esz,
                 sc = (J->smask & 4u) ? 0 : esz;
    size_t i;
    uint32_t acc = 0;
    for (i = J->lo; i < J->hi; i++) {
        uint32_t fl = 0;
        const uint8_t *pa = J->a ? J->a + i * sa : NULL;
        const uint8_t *pb = J->b ? J->b + i * sb : NULL;
        const uint8_t *pc = J->c ? J->c + i * sc : NULL;
        uint8_t tmp[32];
        if (J->op == 256) {
            static const uint8_t zero[32];
            oc_div(J->fmt, J->rnd, pa ? pa : zero, pb ? pb : zero, tmp, &fl);
        } else if (J->op == 257) {
            static const uint8_t zero[32];
            oc_sqrt(J->fmt, J->rnd, pa ? pa : zero, tmp, &fl);
        } else {
            oc_compute(J->fmt, J->op, J->rnd, pa, pb, pc, tmp, &fl);
        }
        memcpy(J->d + i * esz, tmp, esz);
        if (J->lane_flags)
            J->lane_flags[i] = (uint8_t)fl;
        acc |= fl;
    }
    J->flags = acc;
    return NULL;
}

int oc_run(int fmt, int op, int rnd, const void *a, const void *b,
           const void *c, void *d, size_t n, uint32_t scalar_mask,
           int threads, uint32_t *flags_out)
{
    return oc_run_lanes(fmt, op, rnd, a, b, c, d, n, scalar_mask, threads, flags_out, NULL);
}

int oc_run_lanes(int fmt, int op, int rnd, const void *a, const void *b,
                 const void *c, void *d, size_t n, uint32_t scalar_mask,
                 int threads, uint32_t *flags_out, uint8_t *lane_flags)
{
    run_job jobs[256];
    pthread_t tid[256];
    uint32_t acc = 0;
    int t;
    if (!fmt_ok(fmt)) return 1;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    for (t = 0; t < threads; t++) {
        run_job *J = &jobs[t];
        J->fmt = fmt; J->op = op; J->rnd = rnd;
        J->a = (const uint8_t *)a; J->b = (const uint8_t *)b;
        J->c = (const uint8_t *)c; J->d = (uint8_t *)d;
        J->lane_flags = lane_flags;
        J->lo = n * (size_t)t / (size_t)threads;
        J->hi = n * (size_t)(t + 1) / (size_t)threads;
        J->smask = scalar_mask;
        J->flags = 0;
    }
    if (threads == 1) {
        run_slice(&jobs[0]);
    } else {
        for (t = 0; t < threads; t++)
            pthread_create(&tid[t], NULL, run_slice, &jobs[t]);
        for (t = 0; t < threads; t++)
            pthread_join(tid[t], NULL);
    }
    for (t = 0; t < threads; t++)
        acc |= jobs[t].flags;
    if (flags_out)
        *flags_out = acc;
    return 0;
}
