/* The oracle's reductions: cft_reduce and cft_reduce_seg restated.
 *
 * TEST INFRASTRUCTURE ONLY - see sf_oracle.h.  What each part follows:
 *   the tree        host/src/softfloat.c:880-925 (cft_sf_reduce): an empty
 *                   range is +0, one element is its leaf, and a longer
 *                   range splits so its LEFT child is the largest power of
 *                   two strictly inside it; every node is ADD (a + c) at
 *                   the call's rounding attribute
 *   the leaves      host/src/softfloat.c:858-878 (reduce_leaf): SUM's leaf
 *                   is the element, DOT's the rounded product a*b
 *   the composed    host/src/device.c:2759-3062 (reduce_body): n == 0 is
 *   reductions      the identity (+0; -inf for MAXALL); SUMSQ is DOT over
 *                   (a, a) and SUMABS an ABS pass then SUM, each followed
 *                   by 754-2019 9.4's override (device.c:2678-2714: a NaN
 *                   result over an array holding an infinity is +inf, with
 *                   INVALID only when some element is a signalling NaN);
 *                   MAXALL halves with the elementwise MAX, pairing i with
 *                   i + m/2 and carrying an odd last element
 *                   (device.c:2950-3010)
 *   segments        host/src/device.c:3219-3400 (cft_reduce_seg): result s
 *                   is the reduction over slice s of `seg` elements, the
 *                   flags the OR over the slices
 * The code is our own, over the bignum arithmetic of sf_oracle.c.
 */
#include "sf_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { OP_ADD = 1, OP_MUL = 3, OP_ABS = 4, OP_MAX = 8,
       OP_SUM = 24, OP_DOT = 25, OP_SUMSQ = 28, OP_SUMABS = 29, OP_MAXALL = 31 };

static void load_el(const oc_fmt *f, const uint8_t *p, size_t i, oc_big *r)
{
    oc_load(r, p + i * (size_t)(f->width / 8), f->width / 8);
}

/* host/src/softfloat.c:880-925 over [lo, hi) with op SUM or DOT */
static void tree(const oc_fmt *f, int op, int rnd, const uint8_t *a, const uint8_t *b,
                 size_t lo, size_t hi, oc_big *out, uint32_t *flags)
{
    oc_big l, r, z;
    uint32_t lf = 0, rf = 0, af = 0;
    size_t m, k, mid;
    *flags = 0;
    memset(&z, 0, sizeof z);
    if (hi == lo) {
        memset(out, 0, sizeof *out);
        return;
    }
    if (hi - lo == 1) {
        if (op == OP_DOT) {
            oc_big x, y;
            load_el(f, a, lo, &x);
            load_el(f, b, lo, &y);
            oc_compute_big(f, OP_MUL, rnd, &x, &y, &z, out, flags);
        } else {
            load_el(f, a, lo, out);
        }
        return;
    }
    m = hi - lo;
    k = 1;
    while (k < m)
        k <<= 1;
    mid = lo + (k >> 1);
    tree(f, op, rnd, a, b, lo, mid, &l, &lf);
    tree(f, op, rnd, a, b, mid, hi, &r, &rf);
    oc_compute_big(f, OP_ADD, rnd, &l, &z, &r, out, &af);
    *flags = lf | rf | af;
}

static int is_nan(const oc_fmt *f, const uint8_t *p)
{
    const int esz = f->width / 8;
    int i, expo_all = 1, frac = 0;
    /* exponent field: bits [width-2 .. man_w], fraction below */
    for (i = f->man_w; i < f->width - 1; i++)
        if (!((p[i / 8] >> (i % 8)) & 1))
            expo_all = 0;
    for (i = 0; i < f->man_w; i++)
        if ((p[i / 8] >> (i % 8)) & 1)
            frac = 1;
    (void)esz;
    return expo_all && frac;
}

/* the class bits 9.4's override looks at: 1 an infinity, 2 a signalling NaN */
static unsigned special_bits(const oc_fmt *f, const uint8_t *p)
{
    int i, expo_all = 1, frac = 0;
    for (i = f->man_w; i < f->width - 1; i++)
        if (!((p[i / 8] >> (i % 8)) & 1))
            expo_all = 0;
    if (!expo_all)
        return 0;
    for (i = 0; i < f->man_w; i++)
        if ((p[i / 8] >> (i % 8)) & 1)
            frac = 1;
    if (!frac)
        return 1u;
    return ((p[(f->man_w - 1) / 8] >> ((f->man_w - 1) % 8)) & 1) ? 0u : 2u;
}

static void store_inf(const oc_fmt *f, int sign, uint8_t *out)
{
    const int esz = f->width / 8;
    int i;
    memset(out, 0, (size_t)esz);
    for (i = f->man_w; i < f->width - 1; i++)
        out[i / 8] |= (uint8_t)(1u << (i % 8));
    if (sign)
        out[esz - 1] |= 0x80u;
}

int oc_reduce(int fmt, int op, int rnd, const void *a_, const void *b_, void *d_, size_t n,
              uint32_t *flags_out)
{
    const oc_fmt *f;
    const uint8_t *a = (const uint8_t *)a_, *b = (const uint8_t *)b_;
    uint8_t *d = (uint8_t *)d_;
    size_t esz, i;
    uint32_t fl = 0;
    oc_big r;
    if (fmt < 0 || fmt > 3)
        return 1;
    f = &oc_formats[fmt];
    esz = (size_t)f->width / 8;
    if (n == 0) {
        if (op == OP_MAXALL)
            store_inf(f, 1, d);
        else
            memset(d, 0, esz);
        if (flags_out) *flags_out = 0;
        return 0;
    }
    if (op == OP_MAXALL) {
        uint8_t *buf = (uint8_t *)malloc((n + 1) * esz);
        size_t m = n;
        if (!buf)
            return 1;
        memcpy(buf, a, n * esz);
        while (m > 1) {
            const size_t h = m / 2;
            const int odd = (int)(m & 1u);
            for (i = 0; i < h; i++) {
                uint32_t f1 = 0;
                uint8_t t[32];
                oc_compute(fmt, OP_MAX, rnd, buf + i * esz, buf + (i + h) * esz, NULL, t, &f1);
                fl |= f1;
                memcpy(buf + i * esz, t, esz);
            }
            if (odd)
                memmove(buf + h * esz, buf + 2 * h * esz, esz);
            m = h + (size_t)odd;
        }
        memcpy(d, buf, esz);
        free(buf);
        if (flags_out) *flags_out = fl;
        return 0;
    }
    if (op == OP_SUM || op == OP_DOT || op == OP_SUMSQ) {
        const uint8_t *bb = op == OP_SUMSQ ? a : b;
        tree(f, op == OP_SUM ? OP_SUM : OP_DOT, rnd, a, bb, 0, n, &r, &fl);
        oc_store(&r, d, (int)esz);
    } else if (op == OP_SUMABS) {
        uint8_t *mag = (uint8_t *)malloc(n * esz);
        if (!mag)
            return 1;
        memcpy(mag, a, n * esz);
        for (i = 0; i < n; i++)
            mag[i * esz + esz - 1] &= 0x7Fu;     /* 5.5.1 abs: the sign bit, no flags */
        tree(f, OP_SUM, rnd, mag, NULL, 0, n, &r, &fl);
        free(mag);
        oc_store(&r, d, (int)esz);
    } else {
        return 1;
    }
    if ((op == OP_SUMSQ || op == OP_SUMABS) && is_nan(f, d)) {
        unsigned bits = 0;
        for (i = 0; i < n; i++)
            bits |= special_bits(f, a + i * esz);
        if (bits & 1u) {
            store_inf(f, 0, d);
            fl = (bits & 2u) ? OC_FLAG_INVALID : 0u;
        }
    }
    if (flags_out) *flags_out = fl;
    return 0;
}

int oc_reduce_seg(int fmt, int op, int rnd, const void *a_, const void *b_, void *d_,
                  size_t n, size_t seg, uint32_t *flags_out)
{
    const uint8_t *a = (const uint8_t *)a_, *b = (const uint8_t *)b_;
    uint8_t *d = (uint8_t *)d_;
    size_t esz, s;
    uint32_t acc = 0;
    if (fmt < 0 || fmt > 3 || seg == 0 || n % seg)
        return 1;
    esz = (size_t)oc_formats[fmt].width / 8;
    for (s = 0; s < n / seg; s++) {
        uint32_t fl = 0;
        if (oc_reduce(fmt, op, rnd, a + s * seg * esz, b ? b + s * seg * esz : NULL,
                      d + s * esz, seg, &fl))
            return 1;
        acc |= fl;
    }
    if (flags_out) *flags_out = acc;
    return 0;
}
