/* Seeded synthetic operands for the reference's benchmark configurations.
 *
 * Counter-based: element i's operands are a pure function of (seed, i), so
 * any slice, any thread count and any machine produce the same bytes (the
 * property the reference's own differential streams have,
 * host/tools/fx256_stream.h:11-41).  Built into libcftwork.so; the product
 * library does not use it.
 *
 *   cftw_elem_c1   binary128 FMA operands: per operand a random sign; 98%
 *                  an unbiased exponent uniform in [-64, 64] with uniform
 *                  significand bits; 1% an exponent uniform over the whole
 *                  range [emin - p + 1, emax], subnormals included; 1% a
 *                  special drawn from {+-0, +-inf, qNaN, sNaN}.
 *   cftw_elem_c2   binary256 operands for add/mul/fma/div/sqrt: exponents
 *                  uniform over the full normal range, 3% subnormals, 2%
 *                  specials; 5% of elements near a rounding boundary: a, b
 *                  drawn, the exact product P formed, and c set to
 *                  H - (P mod 2u) + k*ulp(c), k in {-1, 0, 1}, where u is
 *                  the result ulp at P's binade and H = u/2 - so the exact
 *                  P + c lies within one ulp of c of a halfway point.
 */
#include <pthread.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    int nl, expw, p, bias;
} wfmt;

static const wfmt WF[4] = {{1, 8, 24, 127}, {2, 11, 53, 1023},
                           {4, 15, 113, 16383}, {8, 19, 237, 262143}};

static uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* the k-th 64-bit draw of (seed, element, operand) */
static uint64_t draw(uint64_t seed, uint64_t i, uint32_t opnd, uint32_t k)
{
    return mix(mix(seed * 0x100000001B3ull + (uint64_t)opnd * 0x1000193ull + k) ^
               (i * 0xD6E8FEB86659FD93ull));
}

/* uniform integer in [0, n) */
static uint64_t below(uint64_t r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)r * n) >> 64);
}

/* fill the fraction field (man_w bits) with random bits; x has nl limbs */
static void rand_frac(const wfmt *f, uint32_t *x, uint64_t seed, uint64_t i,
                      uint32_t opnd, uint32_t k0)
{
    const int manw = 32 * f->nl - 1 - f->expw;
    int l;
    for (l = 0; l < f->nl; l++) {
        const uint64_t r = draw(seed, i, opnd, k0 + (uint32_t)l);
        x[l] = (uint32_t)r;
    }
    x[f->nl - 1] &= (1u << (manw - 32 * (f->nl - 1))) - 1u;
}

static void set_exp_sign(const wfmt *f, uint32_t *x, uint32_t ef, int sign)
{
    const int topf = 32 * f->nl - 1 - f->expw - 32 * (f->nl - 1);
    x[f->nl - 1] |= (ef << topf) | ((uint32_t)sign << 31);
}

static void special(const wfmt *f, uint32_t *x, uint64_t r, uint64_t r2)
{
    const int topf = 32 * f->nl - 1 - f->expw - 32 * (f->nl - 1);
    const uint32_t emask = (1u << f->expw) - 1u;
    const int kind = (int)below(r, 6);           /* +0 -0 +inf -inf qNaN sNaN */
    int l;
    for (l = 0; l < f->nl; l++)
        x[l] = 0;
    switch (kind) {
    case 0: break;
    case 1: x[f->nl - 1] = 0x80000000u; break;
    case 2: x[f->nl - 1] = emask << topf; break;
    case 3: x[f->nl - 1] = (emask << topf) | 0x80000000u; break;
    case 4:
        x[0] = (uint32_t)r2;                     /* any payload, quiet bit set */
        x[f->nl - 1] |= (emask << topf) | (1u << (topf - 1)) |
                        ((uint32_t)(r2 >> 63) << 31);
        break;
    default:
        x[0] = (uint32_t)r2 | 1u;                /* nonzero payload, quiet clear */
        x[f->nl - 1] |= (emask << topf) | ((uint32_t)(r2 >> 63) << 31);
        x[f->nl - 1] &= ~(1u << (topf - 1));
        break;
    }
}

/* A subnormal or normal with its leading bit at unbiased exponent E in
 * [emin - p + 1, emax]; fraction bits below the leading bit random. */
static void at_exponent(const wfmt *f, uint32_t *x, int E, int sign, uint64_t seed,
                        uint64_t i, uint32_t opnd)
{
    const int emin = 1 - f->bias, manw = f->p - 1;
    rand_frac(f, x, seed, i, opnd, 8);
    if (E >= emin) {
        set_exp_sign(f, x, (uint32_t)(E + f->bias), sign);
        return;
    }
    /* subnormal: leading bit at fraction bit manw - (emin - E) */
    {
        const int lead = manw - (emin - E);     /* 0 .. manw-1 */
        int b;
        for (b = lead + 1; b < manw; b++)
            x[b >> 5] &= ~(1u << (b & 31));
        x[lead >> 5] |= 1u << (lead & 31);
        x[f->nl - 1] |= (uint32_t)sign << 31;
    }
}

static void gen_c1(uint64_t seed, uint64_t i, uint32_t opnd, uint32_t *x)
{
    const wfmt *f = &WF[2];
    const uint64_t r0 = draw(seed, i, opnd, 0), r1 = draw(seed, i, opnd, 1);
    const int sign = (int)(r1 >> 63);
    const uint64_t cls = below(r0, 100);
    if (cls < 1) {
        special(f, x, r1, draw(seed, i, opnd, 2));
    } else if (cls < 2) {
        const int lo = (1 - f->bias) - f->p + 1, hi = f->bias;
        at_exponent(f, x, lo + (int)below(draw(seed, i, opnd, 3), (uint64_t)(hi - lo + 1)), sign,
                    seed, i, opnd);
    } else {
        at_exponent(f, x, -64 + (int)below(draw(seed, i, opnd, 3), 129), sign, seed, i, opnd);
    }
}

/* ---- config 2 ---------------------------------------------------------- */

static void gen_c2_operand(uint64_t seed, uint64_t i, uint32_t opnd, uint32_t *x)
{
    const wfmt *f = &WF[3];
    const uint64_t r0 = draw(seed, i, opnd, 0), r1 = draw(seed, i, opnd, 1);
    const int sign = (int)(r1 >> 63);
    const uint64_t cls = below(r0, 100);
    if (cls < 2) {
        special(f, x, r1, draw(seed, i, opnd, 2));
    } else if (cls < 5) {
        /* subnormal: leading bit anywhere in the fraction field */
        const int emin = 1 - f->bias;
        at_exponent(f, x, emin - 1 - (int)below(draw(seed, i, opnd, 3), (uint64_t)(f->p - 1)),
                    sign, seed, i, opnd);
    } else {
        const int emin = 1 - f->bias, emax = f->bias;
        at_exponent(f, x,
                    emin + (int)below(draw(seed, i, opnd, 3), (uint64_t)(emax - emin + 1)),
                    sign, seed, i, opnd);
    }
}

/* 8 x 8 limbs -> 16 */
static void mul8(const uint32_t *a, const uint32_t *b, uint32_t *p)
{
    int i, j;
    memset(p, 0, 16 * sizeof *p);
    for (i = 0; i < 8; i++) {
        uint64_t c = 0;
        for (j = 0; j < 8; j++) {
            c += (uint64_t)a[i] * b[j] + p[i + j];
            p[i + j] = (uint32_t)c;
            c >>= 32;
        }
        p[i + 8] = (uint32_t)c;
    }
}

static int bitlen16(const uint32_t *x)
{
    int i;
    for (i = 15; i >= 0; i--)
        if (x[i])
            return 32 * i + 32 - __builtin_clz(x[i]);
    return 0;
}

static int bit16(const uint32_t *x, int k)
{
    return k < 0 || k >= 512 ? 0 : (int)((x[k >> 5] >> (k & 31)) & 1u);
}

/* The near-boundary FMA case: a, b normal with exponents that keep the
 * product and c well inside the range; c built from P's low bits. */
static void gen_c2_boundary(uint64_t seed, uint64_t i, uint32_t *a, uint32_t *b, uint32_t *c)
{
    const wfmt *f = &WF[3];
    const int p = f->p, manw = p - 1, bias = f->bias;
    uint32_t ma[8], mb[8], P[16], T[16];
    int ea, eb, Lp, k, lsb, j, top, lo;
    const uint64_t r = draw(seed, i, 9, 0);
    rand_frac(f, a, seed, i, 10, 0);
    rand_frac(f, b, seed, i, 11, 0);
    ea = -1000 + (int)below(r, 2001);
    eb = -1000 + (int)below(draw(seed, i, 9, 1), 2001);
    memcpy(ma, a, sizeof ma);
    memcpy(mb, b, sizeof mb);
    ma[7] |= 1u << (manw - 224);
    mb[7] |= 1u << (manw - 224);
    set_exp_sign(f, a, (uint32_t)(ea + bias), (int)(draw(seed, i, 9, 2) >> 63));
    set_exp_sign(f, b, (uint32_t)(eb + bias), (int)(draw(seed, i, 9, 3) >> 63));
    mul8(ma, mb, P);
    Lp = bitlen16(P);                            /* 2p-1 or 2p */
    /* the result ulp at P's binade is bit (Lp - p) of P; H is bit Lp-p-1.
     * T = H - (P mod 2^(Lp-p)), |T| < H: c's magnitude and sign */
    k = (int)below(draw(seed, i, 9, 4), 3) - 1;  /* -1, 0, +1 ulp of c */
    {
        const int ub = Lp - p;                   /* ulp bit */
        uint32_t low[16], h[16];
        int neg = 0;
        memset(low, 0, sizeof low);
        memset(h, 0, sizeof h);
        for (j = 0; j < ub; j++)
            if (bit16(P, j))
                low[j >> 5] |= 1u << (j & 31);
        h[(ub - 1) >> 5] |= 1u << ((ub - 1) & 31);
        /* T = h - low (signed) */
        {
            int64_t br = 0;
            int ge = 1;
            for (j = 15; j >= 0; j--)
                if (h[j] != low[j]) { ge = h[j] > low[j]; break; }
            const uint32_t *x = ge ? h : low, *y = ge ? low : h;
            for (j = 0; j < 16; j++) {
                const int64_t t = (int64_t)x[j] - y[j] - br;
                T[j] = (uint32_t)t;
                br = t < 0;
            }
            neg = !ge;
        }
        /* c = (-1)^(sp ^ neg) * (T + k) * 2^(ep); T's lsb weight is P's */
        {
            int64_t cy = k;
            if (neg) cy = -cy;                   /* k counts toward the sum */
            for (j = 0; j < 16 && cy; j++) {
                const int64_t t = (int64_t)T[j] + cy;
                T[j] = (uint32_t)t;
                cy = t >> 32;
            }
        }
        top = bitlen16(T);
        memset(c, 0, 8 * sizeof *c);
        if (top == 0)
            return;                              /* c = +0: P itself, exact tie */
        /* normalise T to p bits: it has at most p bits (|T| < 2^(ub-1)) */
        lo = top > p ? top - p : 0;              /* bits below lo are zero */
        lsb = lo;
        {
            const int sp = (int)((a[7] ^ b[7]) >> 31);
            const int ep = (ea - manw) + (eb - manw);
            const int E = ep + top - 1;          /* leading-bit exponent of c */
            uint32_t m[8];
            memset(m, 0, sizeof m);
            for (j = 0; j < top - lsb && j < p; j++)
                if (bit16(T, lsb + j))
                    m[j >> 5] |= 1u << (j & 31);
            /* place: leading bit at fraction bit manw (hidden) */
            {
                const int sh = p - (top - lsb);
                uint32_t mm[8];
                memset(mm, 0, sizeof mm);
                for (j = 0; j < p; j++) {
                    const int src = j - sh;
                    if (src >= 0 && ((m[src >> 5] >> (src & 31)) & 1u))
                        mm[j >> 5] |= 1u << (j & 31);
                }
                mm[7] &= (1u << (manw - 224)) - 1u;   /* drop the hidden bit */
                memcpy(c, mm, sizeof mm);
            }
            set_exp_sign(f, c, (uint32_t)(E + bias), sp ^ neg);
        }
    }
}

typedef struct {
    int cfg;
    uint64_t seed, lo, hi;
    uint32_t *a, *b, *c;
} wjob;

static void *wrun(void *arg)
{
    wjob *J = (wjob *)arg;
    uint64_t i;
    for (i = J->lo; i < J->hi; i++) {
        if (J->cfg == 1) {
            gen_c1(J->seed, i, 0, J->a + 4 * i);
            gen_c1(J->seed, i, 1, J->b + 4 * i);
            gen_c1(J->seed, i, 2, J->c + 4 * i);
        } else {
            const int nb = below(draw(J->seed, i, 7, 0), 100) < 5;
            if (nb) {
                gen_c2_boundary(J->seed, i, J->a + 8 * i, J->b + 8 * i, J->c + 8 * i);
            } else {
                gen_c2_operand(J->seed, i, 0, J->a + 8 * i);
                gen_c2_operand(J->seed, i, 1, J->b + 8 * i);
                gen_c2_operand(J->seed, i, 2, J->c + 8 * i);
            }
        }
    }
    return NULL;
}

static int gen(int cfg, uint64_t seed, uint64_t first, size_t n, void *a, void *b, void *c,
               int threads)
{
    wjob jobs[64];
    pthread_t tid[64];
    const int nl = cfg == 1 ? 4 : 8;
    int t;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    for (t = 0; t < threads; t++) {
        jobs[t].cfg = cfg;
        jobs[t].seed = seed;
        jobs[t].lo = first + n * (uint64_t)t / (uint64_t)threads;
        jobs[t].hi = first + n * (uint64_t)(t + 1) / (uint64_t)threads;
        /* the arrays hold elements first .. first+n-1 at their own index 0 */
        jobs[t].a = (uint32_t *)a - (size_t)first * nl;
        jobs[t].b = (uint32_t *)b - (size_t)first * nl;
        jobs[t].c = (uint32_t *)c - (size_t)first * nl;
    }
    if (threads == 1) {
        wrun(&jobs[0]);
        return 0;
    }
    for (t = 0; t < threads; t++)
        pthread_create(&tid[t], NULL, wrun, &jobs[t]);
    for (t = 0; t < threads; t++)
        pthread_join(tid[t], NULL);
    return 0;
}

/* Elements first .. first+n-1 of config 1 (binary128 FMA operands). */
int cftw_elem_c1(uint64_t seed, uint64_t first, size_t n, void *a, void *b, void *c,
                 int threads)
{
    return gen(1, seed, first, n, a, b, c, threads);
}

/* Elements first .. first+n-1 of config 2 (binary256 operands). */
int cftw_elem_c2(uint64_t seed, uint64_t first, size_t n, void *a, void *b, void *c,
                 int threads)
{
    return gen(2, seed, first, n, a, b, c, threads);
}
