/* cft.h - the C ABI of the MI355X engine for the reference library's
 * elementwise binary128/binary256 arithmetic and its sequencer executor.
 *
 * This is a drop-in for loganw234/cft-fp256's `libcft` on that path: the
 * names, enumerator values and struct layouts below are the reference's
 * (host/include/cft.h, ABI 0.18), so a caller compiled against the
 * reference header links and runs against libcft.so from this tree.  Each
 * block cites the reference lines it mirrors.  The text is ours.
 *
 * Device strings for cft_open: "hip:N" opens MI355X device N, N in decimal
 * and nothing else, with the index argument 0 (the grammar of the
 * reference's GPU backend, host/src/backend_hip.cpp:2764-2818).
 * There is no software backend in this library: cft_open(NULL, ...) is
 * CFT_ERR_UNSUPPORTED with a sentence, by design - every result this library
 * returns was computed by the HIP kernels.
 */
#ifndef CFT_H
#define CFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef CFT_API
#define CFT_API __attribute__((visibility("default")))
#endif

/* reference host/include/cft.h:111-129 */
#define CFT_ABI_VERSION_MAJOR 0
#define CFT_ABI_VERSION_MINOR 18
CFT_API uint32_t cft_abi_version(void);  /* major << 16 | minor */

/* reference host/include/cft.h:244-303 */
typedef enum cft_status {
    CFT_OK = 0,
    CFT_ERR_INVALID_ARGUMENT,
    CFT_ERR_UNSUPPORTED,
    CFT_ERR_NO_DEVICE,
    CFT_ERR_ARTIFACT,
    CFT_ERR_BUS_FAULT,
    CFT_ERR_OUT_OF_MEMORY,
    CFT_ERR_TIMEOUT,
    CFT_ERR_INTERNAL
} cft_status;

CFT_API const char *cft_strerror(cft_status s);
/* One sentence about the last failed call, "" after a successful one.
 * Process-global, as the reference's is. */
CFT_API const char *cft_last_error(void);

/* reference host/include/cft.h:312-323 */
typedef enum cft_format {
    CFT_FP32  = 0,
    CFT_FP64  = 1,
    CFT_FP128 = 2,
    CFT_FP256 = 3
} cft_format;

CFT_API size_t      cft_format_size(cft_format f);
CFT_API const char *cft_format_name(cft_format f);

/* reference host/include/cft.h:325-477 */
typedef enum cft_op {
    CFT_FMA = 0, CFT_ADD = 1, CFT_SUB = 2, CFT_MUL = 3,
    CFT_ABS = 4, CFT_NEG = 5, CFT_COPYSIGN = 6,
    CFT_MIN = 7, CFT_MAX = 8, CFT_MINNUM = 9, CFT_MAXNUM = 10,
    CFT_SELECT = 11, CFT_CMPLT = 12, CFT_CMPLE = 13, CFT_CMPEQ = 14,
    CFT_IAND = 16, CFT_IOR = 17, CFT_IXOR = 18, CFT_IADD = 19,
    CFT_ISUB = 20, CFT_ISHL = 21, CFT_ISHR = 22, CFT_ICMPLT = 23,
    CFT_SUM = 24, CFT_DOT = 25,
    CFT_RECIP_SEED = 26, CFT_RSQRT_SEED = 27,
    CFT_SUMSQ = 28, CFT_SUMABS = 29,
    CFT_IMUL = 30,
    CFT_MAXALL = 31
} cft_op;

CFT_API const char *cft_op_name(cft_op op);

/* reference host/include/cft.h:481-501 */
typedef enum cft_round {
    CFT_RNE = 0, CFT_RTZ = 1, CFT_RDN = 2, CFT_RUP = 3, CFT_RMM = 4
} cft_round;

typedef enum cft_exception {
    CFT_FLAG_INVALID   = 1u << 0,
    CFT_FLAG_DIVBYZERO = 1u << 1,
    CFT_FLAG_OVERFLOW  = 1u << 2,
    CFT_FLAG_UNDERFLOW = 1u << 3,
    CFT_FLAG_INEXACT   = 1u << 4
} cft_exception;

/* ---- devices: reference host/include/cft.h:506-579 ------------------- */
typedef struct cft_device cft_device;

CFT_API cft_status cft_open(const char *artifact, int index, cft_device **out);
CFT_API void       cft_close(cft_device *dev);

typedef struct cft_open_args {
    size_t      struct_size;     /* sizeof(cft_open_args) */
    const char *artifact;
    int         index;
    uint32_t    scratch_depth;   /* 0 (256), or a power of two <= 2048 */
} cft_open_args;

CFT_API cft_status cft_open_ex(const cft_open_args *args, cft_device **out);

/* reference host/include/cft.h:655-772 */
typedef struct cft_caps {
    size_t   struct_size;
    uint32_t format_mask;
    uint32_t tiles;
    uint32_t abi_version;
    uint32_t device_version;
    int      flags_readable;
    char     backend[32];
    uint32_t max_deposits;
    uint32_t max_insns;
    uint32_t max_consts;
    uint32_t seq_features;
    uint32_t max_scratch;
    int      buffers_resident;
} cft_caps;

/* reference host/include/cft.h:785-1042 */
#define CFT_SEQ_FEAT_WIDE_CONST     0x01u
#define CFT_SEQ_FEAT_REGS32         0x02u
#define CFT_SEQ_FEAT_BANK_PTR       0x04u
#define CFT_SEQ_FEAT_KX9            0x08u
#define CFT_ALU_EXT_IMUL            0x10u
#define CFT_SEQ_FEAT_SCRATCH        0x100u
#define CFT_SEQ_FEAT_SCRATCH_IO     0x200u
#define CFT_SEQ_FEAT_SCRATCH_STRICT 0x400u
#define CFT_SEQ_FEAT_SCALAR         0x800u
#define CFT_FEAT_REDUCE_SEG         0x1000u
#define CFT_SEQ_FEAT_INDEXED        0x2000u
#define CFT_SEQ_FEAT_LANE_MASK      0x4000u
#define CFT_SEQ_FEAT_AUGADD         0x8000u
#define CFT_SEQ_FEAT_SCRATCH_STEP   0x10000u
#define CFT_SEQ_FEAT_LANE_FLAGS     0x20000u
#define CFT_SEQ_FEAT_FLAG_CONTROL   0x40000u
#define CFT_IDX_NONE 0xFFFFFFFFu

CFT_API cft_status cft_get_caps(cft_device *dev, cft_caps *out);
CFT_API int        cft_supports(cft_device *dev, cft_op op, cft_format fmt);

/* ---- elementwise: reference host/include/cft.h:1198-1209 -------------- */
CFT_API cft_status cft_run(cft_device *dev, cft_op op, cft_format fmt,
                           cft_round rnd, const void *a, const void *b,
                           const void *c, void *d, size_t n,
                           uint32_t *flags_out, uint32_t *bus_out);

/* ---- reductions: reference host/include/cft.h:1210-1340 ---------------
 * d receives one element (cft_reduce) or n / seg elements (cft_reduce_seg):
 * the fixed tree whose left child is the largest power of two strictly
 * inside the range, at the call's rounding attribute at every node.
 * Opcodes CFT_SUM, CFT_DOT, CFT_SUMSQ, CFT_SUMABS, CFT_MAXALL. */
CFT_API cft_status cft_reduce(cft_device *dev, cft_op op, cft_format fmt,
                              cft_round rnd, const void *a, const void *b,
                              void *d, size_t n, uint32_t *flags_out,
                              uint32_t *bus_out);
CFT_API cft_status cft_reduce_seg(cft_device *dev, cft_op op, cft_format fmt,
                                  cft_round rnd, const void *a, const void *b,
                                  void *d, size_t n, size_t seg,
                                  uint32_t *flags_out, uint32_t *bus_out);

/* reference host/include/cft.h:1385-1402 */
CFT_API cft_status cft_div(cft_device *dev, cft_format fmt, cft_round rnd,
                           const void *a, const void *b, void *d, size_t n,
                           uint32_t *flags_out, uint32_t *bus_out);
CFT_API cft_status cft_sqrt(cft_device *dev, cft_format fmt, cft_round rnd,
                            const void *a, void *d, size_t n,
                            uint32_t *flags_out, uint32_t *bus_out);

/* ---- resident buffers: reference host/include/cft.h:2653-2704 --------- */
typedef struct cft_buffer cft_buffer;

CFT_API cft_status cft_alloc(cft_device *dev, size_t bytes, cft_buffer **out);
CFT_API void      *cft_buffer_data(cft_buffer *buf);
CFT_API cft_status cft_buffer_to_device(cft_buffer *buf);
CFT_API cft_status cft_buffer_from_device(cft_buffer *buf);
CFT_API void       cft_buffer_free(cft_buffer *buf);

typedef struct cft_buffer_info {
    size_t   struct_size;
    size_t   bytes;
    int      resident;
    int      device_authority;
    uint64_t resident_binds;
    uint64_t staged_binds;
    char     staged_why[112];
} cft_buffer_info;

CFT_API cft_status cft_buffer_get_info(cft_buffer *buf, cft_buffer_info *out);

/* ---- programs: reference host/include/cft.h:2727-3052 ------------------ */
typedef struct cft_program cft_program;

#define CFT_STATUS_DEPOSIT_OVERFLOW (1u << 4)
#define CFT_STATUS_SCRATCH_RANGE    (1u << 5)
#define CFT_STATUS_MARKED           (1u << 6)
#define CFT_LANE_DEPOSIT_OVERFLOW   (1u << 5)
#define CFT_LANE_SCRATCH_RANGE      (1u << 6)
#define CFT_LANE_MARKED             (1u << 7)

CFT_API cft_status cft_program_load(cft_device *dev, const void *image,
                                    size_t bytes, cft_program **out);
CFT_API void       cft_program_free(cft_program *prog);

#define CFT_PROG_FLAG_BANK_EXT       (1u << 0)
#define CFT_PROG_FLAG_SCRATCH_IO     (1u << 1)
#define CFT_PROG_FLAG_SCRATCH_STRICT (1u << 2)

typedef struct cft_program_info {
    size_t     struct_size;
    cft_format format;
    uint32_t   max_deposits;
    uint32_t   n_insns;
    uint32_t   n_consts;
    uint32_t   flags;
    uint32_t   n_scratch_in;
    uint32_t   n_scratch_out;
    uint32_t   scratch_used;
} cft_program_info;

CFT_API cft_status cft_program_get_info(cft_program *prog,
                                        cft_program_info *out);

CFT_API cft_status cft_program_run(cft_program *prog, const void *a,
                                   const void *b, const void *c,
                                   void *deposits, uint32_t *counts, size_t n,
                                   uint32_t *flags, uint32_t *bus);
CFT_API cft_status cft_program_run_bank(cft_program *prog, const void *bank,
                                        size_t bank_bytes, const void *a,
                                        const void *b, const void *c,
                                        void *deposits, uint32_t *counts,
                                        size_t n, uint32_t *flags_out,
                                        uint32_t *bus_out);

typedef struct cft_run_args {
    size_t      struct_size;
    const void *a, *b, *c;
    size_t      n;
    const void *bank;        size_t bank_bytes;
    const void *scratch_in;  size_t scratch_in_bytes;
    void       *scratch_out; size_t scratch_out_bytes;
    void       *deposits;    uint32_t *counts;
    uint32_t   *flags_out;   uint32_t *bus_out;
    const uint32_t *idx_a, *idx_b, *idx_c;
    size_t idx_a_src, idx_b_src, idx_c_src;
    const uint32_t *idx_scratch_in;
    size_t idx_scratch_src;
    const uint8_t *lane_mask;
    size_t lane_mask_bytes;
    uint8_t *lane_flags;
    size_t lane_flags_bytes;
} cft_run_args;

CFT_API cft_status cft_program_run_ex(cft_program *prog,
                                      const cft_run_args *args);

/* SHA-256 over the image bytes as loaded, then the bank's (NULL and 0 for a
 * program that carries its own constants): reference host/include/cft.h:
 * 3154-3160, host/src/program.c:2425-2454. */
CFT_API cft_status cft_program_digest(cft_program *prog, const void *bank,
                                      size_t bank_bytes, uint8_t out[32]);

/* reference host/include/cft.h:3122-3135 */
typedef struct cft_elem_args {
    size_t      struct_size;
    const void *a, *b, *c;
    void       *d;
    size_t      n;
    uint32_t    scalar_mask;
    uint32_t   *flags_out;
    uint32_t   *bus_out;
    const uint32_t *idx_a, *idx_b, *idx_c;
    size_t idx_a_src, idx_b_src, idx_c_src;
} cft_elem_args;

CFT_API cft_status cft_run_ex(cft_device *dev, cft_op op, cft_format fmt,
                              cft_round rnd, const cft_elem_args *args);

/* ---- the status word: reference host/include/cft.h:3224-3267 ---------- */
#define CFT_FLAGS_ALL ((uint32_t)(CFT_FLAG_INVALID | CFT_FLAG_DIVBYZERO | \
                                  CFT_FLAG_OVERFLOW | CFT_FLAG_UNDERFLOW | \
                                  CFT_FLAG_INEXACT))

CFT_API void     cft_lower_flags(cft_device *dev, uint32_t mask);
CFT_API void     cft_raise_flags(cft_device *dev, uint32_t mask);
CFT_API int      cft_test_flags(cft_device *dev, uint32_t mask);
CFT_API uint32_t cft_save_all_flags(cft_device *dev);
CFT_API void     cft_restore_flags(cft_device *dev, uint32_t saved,
                                   uint32_t mask);
CFT_API int      cft_test_saved_flags(uint32_t saved, uint32_t mask);

/* ---- entry points off this path ----------------------------------------
 *
 * The reference's Python binding (bindings/python/cftmpfr/_lib.py:181-376,
 * `_bind`) declares every entry point below when it loads a library, so
 * they are exported here with the reference's signatures; the operations
 * themselves (transcendentals, the integer-exponent powers, augmented
 * arithmetic, scaled products, formatOf, character conversions, NaN
 * payloads, the 9.6 magnitude forms) are not on this library's path.  Each
 * clears the error slot, sets one sentence naming itself and returns
 * CFT_ERR_UNSUPPORTED, writing nothing.  A caller that needs them keeps the
 * reference's software library for those calls. */

/* reference host/include/cft.h:1648-1671 (one operand) */
#define CFT_OFFPATH_UNARY(name) \
    CFT_API cft_status name(cft_device *dev, cft_format fmt, cft_round rnd, \
                            const void *a, void *d, size_t n, uint32_t *flags_out);
CFT_OFFPATH_UNARY(cft_exp) CFT_OFFPATH_UNARY(cft_expm1) CFT_OFFPATH_UNARY(cft_exp2)
CFT_OFFPATH_UNARY(cft_log) CFT_OFFPATH_UNARY(cft_log1p) CFT_OFFPATH_UNARY(cft_log2)
CFT_OFFPATH_UNARY(cft_log10) CFT_OFFPATH_UNARY(cft_sinpi) CFT_OFFPATH_UNARY(cft_cospi)
CFT_OFFPATH_UNARY(cft_tanpi) CFT_OFFPATH_UNARY(cft_asin) CFT_OFFPATH_UNARY(cft_acos)
CFT_OFFPATH_UNARY(cft_atan) CFT_OFFPATH_UNARY(cft_asinpi) CFT_OFFPATH_UNARY(cft_acospi)
CFT_OFFPATH_UNARY(cft_atanpi) CFT_OFFPATH_UNARY(cft_sin) CFT_OFFPATH_UNARY(cft_cos)
CFT_OFFPATH_UNARY(cft_tan) CFT_OFFPATH_UNARY(cft_sinh) CFT_OFFPATH_UNARY(cft_cosh)
CFT_OFFPATH_UNARY(cft_tanh) CFT_OFFPATH_UNARY(cft_asinh) CFT_OFFPATH_UNARY(cft_acosh)
CFT_OFFPATH_UNARY(cft_atanh) CFT_OFFPATH_UNARY(cft_exp2m1) CFT_OFFPATH_UNARY(cft_exp10)
CFT_OFFPATH_UNARY(cft_exp10m1) CFT_OFFPATH_UNARY(cft_log2p1) CFT_OFFPATH_UNARY(cft_log10p1)
CFT_OFFPATH_UNARY(cft_rsqrt)
#undef CFT_OFFPATH_UNARY

/* reference host/include/cft.h:1672-1690 (two operands) */
#define CFT_OFFPATH_BINARY(name) \
    CFT_API cft_status name(cft_device *dev, cft_format fmt, cft_round rnd, \
                            const void *a, const void *b, void *d, size_t n, \
                            uint32_t *flags_out);
CFT_OFFPATH_BINARY(cft_pow) CFT_OFFPATH_BINARY(cft_hypot) CFT_OFFPATH_BINARY(cft_atan2)
CFT_OFFPATH_BINARY(cft_atan2pi) CFT_OFFPATH_BINARY(cft_powr)
#undef CFT_OFFPATH_BINARY

/* reference host/include/cft.h:2119-2130 (an integer exponent array) */
CFT_API cft_status cft_pown(cft_device *dev, cft_format fmt, cft_round rnd,
                            const void *a, const int64_t *n, void *d,
                            size_t count, uint32_t *flags_out);
CFT_API cft_status cft_compound(cft_device *dev, cft_format fmt, cft_round rnd,
                                const void *a, const int64_t *n, void *d,
                                size_t count, uint32_t *flags_out);
CFT_API cft_status cft_rootn(cft_device *dev, cft_format fmt, cft_round rnd,
                             const void *a, const int64_t *n, void *d,
                             size_t count, uint32_t *flags_out);

/* reference host/include/cft.h:2335-2348 (9.4 scaled products) */
CFT_API cft_status cft_scaled_prod(cft_device *dev, cft_format fmt, cft_round rnd,
                                   const void *a, void *pr, int64_t *scale_out,
                                   size_t n, uint32_t *flags_out);
CFT_API cft_status cft_scaled_prod_sum(cft_device *dev, cft_format fmt, cft_round rnd,
                                       const void *a, const void *b, void *pr,
                                       int64_t *scale_out, size_t n, uint32_t *flags_out);
CFT_API cft_status cft_scaled_prod_diff(cft_device *dev, cft_format fmt, cft_round rnd,
                                        const void *a, const void *b, void *pr,
                                        int64_t *scale_out, size_t n, uint32_t *flags_out);

/* reference host/include/cft.h:2440-2451 (9.5 augmented arithmetic) */
CFT_API cft_status cft_augmented_add(cft_device *dev, cft_format fmt, const void *a,
                                     const void *b, void *r, void *e, size_t n,
                                     uint32_t *flags_out);
CFT_API cft_status cft_augmented_sub(cft_device *dev, cft_format fmt, const void *a,
                                     const void *b, void *r, void *e, size_t n,
                                     uint32_t *flags_out);
CFT_API cft_status cft_augmented_mul(cft_device *dev, cft_format fmt, const void *a,
                                     const void *b, void *r, void *e, size_t n,
                                     uint32_t *flags_out);

/* reference host/include/cft.h:3511-3541 (5.4.1 formatOf) */
#define CFT_OFFPATH_FORMATOF2(name) \
    CFT_API cft_status name(cft_device *dev, cft_format sfmt, cft_format dfmt, \
                            cft_round rnd, const void *a, const void *b, void *d, \
                            size_t n, uint32_t *flags_out, uint32_t *bus_out);
CFT_OFFPATH_FORMATOF2(cft_formatof_add) CFT_OFFPATH_FORMATOF2(cft_formatof_sub)
CFT_OFFPATH_FORMATOF2(cft_formatof_mul) CFT_OFFPATH_FORMATOF2(cft_formatof_div)
#undef CFT_OFFPATH_FORMATOF2
CFT_API cft_status cft_formatof_sqrt(cft_device *dev, cft_format sfmt, cft_format dfmt,
                                     cft_round rnd, const void *a, void *d, size_t n,
                                     uint32_t *flags_out, uint32_t *bus_out);
CFT_API cft_status cft_formatof_fma(cft_device *dev, cft_format sfmt, cft_format dfmt,
                                    cft_round rnd, const void *a, const void *b,
                                    const void *c, void *d, size_t n, uint32_t *flags_out,
                                    uint32_t *bus_out);

/* reference host/include/cft.h:1899-1964 (5.12 character conversions).
 * cft_format_decimal_digits is a formula of the format and is computed:
 * 1 + ceil(p * log10 2), 9, 17, 36 and 73; 0 for a format out of range. */
CFT_API size_t cft_format_decimal_digits(cft_format fmt);
CFT_API cft_status cft_from_decimal_char(cft_device *dev, cft_format fmt, cft_round rnd,
                                         const char *const *in, void *d, size_t n,
                                         size_t *bad_index, uint32_t *flags_out);
CFT_API cft_status cft_to_decimal_char(cft_device *dev, cft_format fmt, cft_round rnd,
                                       const void *a, size_t digits, char *out, size_t cap,
                                       size_t *len, uint32_t *flags_out);
CFT_API cft_status cft_from_hex_char(cft_device *dev, cft_format fmt, cft_round rnd,
                                     const char *const *in, void *d, size_t n,
                                     size_t *bad_index, uint32_t *flags_out);
CFT_API cft_status cft_to_hex_char(cft_device *dev, cft_format fmt, const void *a, char *out,
                                   size_t cap, size_t *len);

/* reference host/include/cft.h:1996-2002 (9.7 NaN payloads) */
CFT_API cft_status cft_get_payload(cft_device *dev, cft_format fmt, const void *a, void *d,
                                   size_t n);
CFT_API cft_status cft_set_payload(cft_device *dev, cft_format fmt, const void *a, void *d,
                                   size_t n);
CFT_API cft_status cft_set_payload_signaling(cft_device *dev, cft_format fmt, const void *a,
                                             void *d, size_t n);

/* reference host/include/cft.h:3360-3371 (9.6 magnitude forms) */
#define CFT_OFFPATH_MAG(name) \
    CFT_API cft_status name(cft_device *dev, cft_format fmt, const void *a, const void *b, \
                            void *d, size_t n, uint32_t *flags_out);
CFT_OFFPATH_MAG(cft_min_mag) CFT_OFFPATH_MAG(cft_max_mag)
CFT_OFFPATH_MAG(cft_minnum_mag) CFT_OFFPATH_MAG(cft_maxnum_mag)
#undef CFT_OFFPATH_MAG

/* reference host/include/cft.h:3307-3309 (5.7.1).  All three are 0 here:
 * this library carries a subset of 754's required operations (none of the
 * character conversions, for one), so it claims no version. */
CFT_API int cft_is754version1985(void);
CFT_API int cft_is754version2008(void);
CFT_API int cft_is754version2019(void);

#ifdef __cplusplus
}
#endif

#endif /* CFT_H */
