/* The oracle: a CPU restatement of the reference's arithmetic contract.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker.
 * The product (libcft.so) never links, loads or calls it.
 *
 * What it restates, with the reference file each part follows:
 *   formats            host/src/softfloat.c:10-16
 *   unpack             host/src/softfloat.c:85-116
 *   rounding           host/src/softfloat.c:120-261 (tininess after rounding)
 *   fma + steering     host/src/softfloat.c:324-468
 *   other ALU ops      host/src/softfloat.c:470-787, 939-1009
 *   div, sqrt          the correctly-rounded contract of
 *                      python/cft_golden/softfloat.py:440-506 (read as text)
 *   elementwise call   host/src/device.c:1699-2174 (NULL operand = +0,
 *                      scalar operand = element 0, flags OR)
 *   reductions         host/src/softfloat.c:858-925, host/src/device.c:
 *                      2678-3062 and 3219-3400 (rd_oracle.c)
 *
 * The arithmetic is our own: exact integer arithmetic on a fixed-size
 * little-endian bignum, with an alignment bound that substitutes a far
 * smaller term by a representative in the same rounding interval.
 */
#ifndef CFT_SF_ORACLE_H
#define CFT_SF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#define OC_FLAG_INVALID   1u
#define OC_FLAG_DIVZERO   2u
#define OC_FLAG_OVERFLOW  4u
#define OC_FLAG_UNDERFLOW 8u
#define OC_FLAG_INEXACT   16u

/* 64 x 32 = 2048 bits: the widest exact intermediate is the near-case fma
 * sum, under 4p + 16 bits (about 970 for fp256), and the div numerator
 * (about 2p + 8).  The margin costs nothing: every op runs over `n`. */
#define OC_LIMBS 64

typedef struct {
    int      n;                 /* significant limbs, w[n-1] != 0 or n == 0 */
    uint32_t w[OC_LIMBS];
} oc_big;

typedef struct {
    int     exp_w, man_w, width, prec;
    int32_t bias, emax, emin;
} oc_fmt;

extern const oc_fmt oc_formats[4];

/* One element of `op` on encodings of `fmt` (0 fp32 .. 3 fp256).  Any of
 * a, b, c may be NULL, which reads as +0.  Returns 0. */
int oc_compute(int fmt, int op, int rnd, const uint8_t *a, const uint8_t *b,
               const uint8_t *c, uint8_t *out, uint32_t *flags);

/* Correctly-rounded division and square root of one element. */
int oc_div(int fmt, int rnd, const uint8_t *a, const uint8_t *b, uint8_t *out,
           uint32_t *flags);
int oc_sqrt(int fmt, int rnd, const uint8_t *a, uint8_t *out, uint32_t *flags);

/* Elementwise over n elements, as cft_run computes it: operand element i
 * is a[i] (or a[0] where bit 0 of scalar_mask is set, likewise b and c),
 * NULL reads as +0, flags are the OR.  op 256 is div (a, b), op 257 sqrt (a).
 * Work is split over `threads` POSIX threads (0 or 1: the calling thread). */
int oc_run(int fmt, int op, int rnd, const void *a, const void *b,
           const void *c, void *d, size_t n, uint32_t scalar_mask,
           int threads, uint32_t *flags_out);

/* oc_run, and element i's own flags into lane_flags[i] (n bytes) where
 * lane_flags is not NULL. */
int oc_run_lanes(int fmt, int op, int rnd, const void *a, const void *b,
                 const void *c, void *d, size_t n, uint32_t scalar_mask,
                 int threads, uint32_t *flags_out, uint8_t *lane_flags);

/* The reductions (rd_oracle.c): cft_reduce's one result of op (24 SUM,
 * 25 DOT, 28 SUMSQ, 29 SUMABS, 31 MAXALL) over n elements, and
 * cft_reduce_seg's n / seg results.  b is read by DOT only.  Returns 1 for
 * an argument the entry points refuse. */
int oc_reduce(int fmt, int op, int rnd, const void *a, const void *b, void *d, size_t n,
              uint32_t *flags_out);
int oc_reduce_seg(int fmt, int op, int rnd, const void *a, const void *b, void *d,
                  size_t n, size_t seg, uint32_t *flags_out);

/* Rounding a caller-built exact value: (-1)^sign * m * 2^e, m given as
 * little-endian bytes (mbytes of them).  Used by the tests' closed-form
 * checks.  Returns 1 for m == 0. */
int oc_round_pack(int fmt, int sign, const uint8_t *m, int mbytes, int e,
                  int rnd, uint8_t *out, uint32_t *flags);

/* shared with seq_oracle.c: one element on bignum registers, and the
 * bignum's byte I/O */
void oc_compute_big(const oc_fmt *f, int op, int rnd, const oc_big *a,
                    const oc_big *b, const oc_big *c, oc_big *out,
                    uint32_t *flags);
void oc_load(oc_big *r, const uint8_t *le, int nbytes);
void oc_store(const oc_big *a, uint8_t *le, int nbytes);

#endif
