// Hazard build of the device limb arithmetic: the CPU build of
// cft-fp256_amd/csrc/fx_*.hpp (fx_emu.cpp, included here) against the oracle
// restatement (oracle/sf_oracle.c, linked in) on seeded random operands,
// element by element, bits and flags.  tests/test_emu_hazards.py builds it
// at -O0 and under AddressSanitizer + UndefinedBehaviorSanitizer (host code
// only; nothing here runs on a GPU) and runs it: any mismatch, any
// out-of-bounds access and any undefined operation (a shift by 32 or more,
// signed overflow) fails the run.  Test infrastructure only.
//
//   emu_check <cases per format, operation and mode> <seed>
#include "fx_emu.cpp"

#include <stdio.h>
#include <stdlib.h>

extern "C" {
int oc_compute(int fmt, int op, int rnd, const uint8_t *a, const uint8_t *b, const uint8_t *c,
               uint8_t *out, uint32_t *flags);
int oc_div(int fmt, int rnd, const uint8_t *a, const uint8_t *b, uint8_t *out,
           uint32_t *flags);
int oc_sqrt(int fmt, int rnd, const uint8_t *a, uint8_t *out, uint32_t *flags);
}

namespace {

uint64_t g_state;

uint64_t next()
{
    // xorshift64*
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return g_state * 0x2545F4914F6CDD1DULL;
}

const int kExp[4] = {8, 11, 15, 19};

// One encoding of format fmt (nl 32-bit limbs, little end first): a tenth
// specials, a tenth subnormals, 40% within 8 binades of 1, a fifth with a
// random exponent, a fifth random bits.
void draw(int fmt, uint32_t *x)
{
    const int nl = 1 << fmt, ew = kExp[fmt], fb = 31 - ew;   // fraction bits in the top limb
    const uint32_t bias = (1u << (ew - 1)) - 1, emax = (1u << ew) - 1;
    for (int k = 0; k < nl; k++)
        x[k] = (uint32_t)next();
    const uint32_t sign = (uint32_t)(next() & 1) << 31;
    const uint32_t frac = x[nl - 1] & ((1u << fb) - 1);
    const unsigned r = (unsigned)(next() % 10);
    uint32_t e;
    if (r == 0) {                       // specials
        const unsigned s = (unsigned)(next() % 4);
        for (int k = 0; k < nl - 1; k++)
            x[k] = 0;
        if (s == 0) { x[nl - 1] = sign; return; }                           // +-0
        if (s == 1) { x[nl - 1] = sign | (emax << fb); return; }            // +-inf
        if (s == 2) { x[nl - 1] = (emax << fb) | (1u << (fb - 1)); return; }  // qNaN
        x[nl - 1] = (emax << fb);                                           // sNaN
        x[0] |= 1;
        return;
    }
    if (r == 1) {                       // subnormal: a random-length fraction
        const int keep = (int)(next() % (unsigned)(32 * nl - 1 - ew));
        for (int k = 0; k < nl; k++) {
            const int lo = 32 * k;
            if (lo >= keep) x[k] = 0;
            else if (lo + 32 > keep) x[k] &= (1u << (keep - lo)) - 1;
        }
        x[nl - 1] = (x[nl - 1] & ((1u << fb) - 1)) | sign;
        return;
    }
    if (r < 6)
        e = bias - 8 + (uint32_t)(next() % 17);
    else if (r < 8)
        e = 1 + (uint32_t)(next() % (emax - 1));
    else
        return;                          // random bits as drawn
    x[nl - 1] = sign | (e << fb) | frac;
}

template <int NL>
long check(int fmt, int op, int rnd, long cases)
{
    long bad = 0;
    for (long i = 0; i < cases; i++) {
        uint32_t a[8], b[8], c[8], d0[8], d1[8];
        draw(fmt, a);
        draw(fmt, b);
        draw(fmt, c);
        if (op == 257)
            a[NL - 1] &= 0x7FFFFFFFu;    // sqrt of |a|, as the workload draws it
        uint32_t f0 = 0, f1 = 0;
        one<NL>(op, rnd, a, b, c, d0, &f0);
        if (op == 256)
            oc_div(fmt, rnd, (const uint8_t *)a, (const uint8_t *)b, (uint8_t *)d1, &f1);
        else if (op == 257)
            oc_sqrt(fmt, rnd, (const uint8_t *)a, (uint8_t *)d1, &f1);
        else if (op == 258 || op == 259)
            oc_compute(fmt, op == 258 ? 1 : 2, rnd, (const uint8_t *)a, nullptr,
                       (const uint8_t *)c, (uint8_t *)d1, &f1);
        else
            oc_compute(fmt, op, rnd, (const uint8_t *)a, (const uint8_t *)b,
                       (const uint8_t *)c, (uint8_t *)d1, &f1);
        if (memcmp(d0, d1, 4 * NL) != 0 || f0 != f1) {
            if (bad < 3)
                fprintf(stderr, "mismatch: fmt %d op %d rnd %d case %ld\n", fmt, op, rnd, i);
            bad++;
        }
    }
    return bad;
}

} // namespace

int main(int argc, char **argv)
{
    const long cases = argc > 1 ? atol(argv[1]) : 2000;
    g_state = argc > 2 ? strtoull(argv[2], nullptr, 10) | 1 : 88172645463325252ULL;
    // FMA, ADD, SUB, MUL, DIV, SQRT, the two-term ADD and SUB paths, and the
    // other opcodes (sign, min/max, compare, select, integer, seeds)
    const int ops[] = {0, 1, 2, 3, 256, 257, 258, 259, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14,
                       16, 17, 18, 19, 20, 21, 22, 23, 26, 27, 30, 77};
    long bad = 0, total = 0;
    for (int fmt = 0; fmt < 4; fmt++)
        for (int op : ops)
            for (int rnd = 0; rnd < 5; rnd++) {
                const long n = op < 4 || op >= 256 ? cases : cases / 8 + 1;
                switch (fmt) {
                case 0: bad += check<1>(fmt, op, rnd, n); break;
                case 1: bad += check<2>(fmt, op, rnd, n); break;
                case 2: bad += check<4>(fmt, op, rnd, n); break;
                default: bad += check<8>(fmt, op, rnd, n); break;
                }
                total += n;
            }
    printf("emu_check: %ld cases, %ld mismatches\n", total, bad);
    return bad ? 1 : 0;
}
