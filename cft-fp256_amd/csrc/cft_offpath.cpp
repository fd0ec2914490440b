// The reference ABI's entry points that are not on this library's path.
//
// The reference's Python binding declares them all when it loads a library
// (bindings/python/cftmpfr/_lib.py:181-376), so a library without them
// cannot be loaded by it at all.  Each is exported with the reference's
// signature (include/cft.h) and refuses: the error slot gets one sentence
// naming the call and CFT_ERR_UNSUPPORTED is returned, nothing written.
// cft_format_decimal_digits and the three 5.7.1 predicates are answered.
#include "cft_internal.hpp"

namespace {

cft_status off_path(const char *who, const char *what)
{
    cft_clear_error();
    cft_set_error("%s: %s is not on this library's path (elementwise arithmetic, reductions "
                  "and the sequencer on MI355X); use the reference's software library for it",
                  who, what);
    return CFT_ERR_UNSUPPORTED;
}

} // namespace

#define UNARY(name)                                                                         \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, cft_round, const void *,   \
                                       void *, size_t, uint32_t *)                          \
    {                                                                                       \
        return off_path(#name, "this transcendental function");                            \
    }
UNARY(cft_exp) UNARY(cft_expm1) UNARY(cft_exp2) UNARY(cft_log) UNARY(cft_log1p)
UNARY(cft_log2) UNARY(cft_log10) UNARY(cft_sinpi) UNARY(cft_cospi) UNARY(cft_tanpi)
UNARY(cft_asin) UNARY(cft_acos) UNARY(cft_atan) UNARY(cft_asinpi) UNARY(cft_acospi)
UNARY(cft_atanpi) UNARY(cft_sin) UNARY(cft_cos) UNARY(cft_tan) UNARY(cft_sinh)
UNARY(cft_cosh) UNARY(cft_tanh) UNARY(cft_asinh) UNARY(cft_acosh) UNARY(cft_atanh)
UNARY(cft_exp2m1) UNARY(cft_exp10) UNARY(cft_exp10m1) UNARY(cft_log2p1) UNARY(cft_log10p1)
UNARY(cft_rsqrt)
#undef UNARY

#define BINARY(name)                                                                        \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, cft_round, const void *,   \
                                       const void *, void *, size_t, uint32_t *)            \
    {                                                                                       \
        return off_path(#name, "this transcendental function");                            \
    }
BINARY(cft_pow) BINARY(cft_hypot) BINARY(cft_atan2) BINARY(cft_atan2pi) BINARY(cft_powr)
#undef BINARY

#define INTEXP(name)                                                                        \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, cft_round, const void *,   \
                                       const int64_t *, void *, size_t, uint32_t *)         \
    {                                                                                       \
        return off_path(#name, "this integer-exponent function");                          \
    }
INTEXP(cft_pown) INTEXP(cft_compound) INTEXP(cft_rootn)
#undef INTEXP

extern "C" CFT_API cft_status cft_scaled_prod(cft_device *, cft_format, cft_round,
                                              const void *, void *, int64_t *, size_t,
                                              uint32_t *)
{
    return off_path("cft_scaled_prod", "the 9.4 scaled product");
}

#define SCALED2(name)                                                                       \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, cft_round, const void *,   \
                                       const void *, void *, int64_t *, size_t, uint32_t *) \
    {                                                                                       \
        return off_path(#name, "the 9.4 scaled product");                                  \
    }
SCALED2(cft_scaled_prod_sum) SCALED2(cft_scaled_prod_diff)
#undef SCALED2

#define AUGMENTED(name)                                                                     \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, const void *, const void *, \
                                       void *, void *, size_t, uint32_t *)                  \
    {                                                                                       \
        return off_path(#name, "9.5 augmented arithmetic");                                \
    }
AUGMENTED(cft_augmented_add) AUGMENTED(cft_augmented_sub) AUGMENTED(cft_augmented_mul)
#undef AUGMENTED

#define FORMATOF2(name)                                                                     \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, cft_format, cft_round,     \
                                       const void *, const void *, void *, size_t,          \
                                       uint32_t *, uint32_t *)                              \
    {                                                                                       \
        return off_path(#name, "formatOf arithmetic (two formats a call)");                \
    }
FORMATOF2(cft_formatof_add) FORMATOF2(cft_formatof_sub) FORMATOF2(cft_formatof_mul)
FORMATOF2(cft_formatof_div)
#undef FORMATOF2

extern "C" CFT_API cft_status cft_formatof_sqrt(cft_device *, cft_format, cft_format,
                                                cft_round, const void *, void *, size_t,
                                                uint32_t *, uint32_t *)
{
    return off_path("cft_formatof_sqrt", "formatOf arithmetic (two formats a call)");
}

extern "C" CFT_API cft_status cft_formatof_fma(cft_device *, cft_format, cft_format,
                                               cft_round, const void *, const void *,
                                               const void *, void *, size_t, uint32_t *,
                                               uint32_t *)
{
    return off_path("cft_formatof_fma", "formatOf arithmetic (two formats a call)");
}

// 5.12.2's Pmin: 1 + ceil(p * log10 2) significant digits for precision p
// (24, 53, 113, 237; reference host/include/cft.h:1889-1899).  Since 2^p is
// never a power of ten, ceil(p * log10 2) is the number of decimal digits of
// 2^p, counted exactly here: 9, 17, 36 and 73.
extern "C" CFT_API size_t cft_format_decimal_digits(cft_format fmt)
{
    static const int kPrec[4] = {24, 53, 113, 237};
    const int f = (int)fmt;
    if (f < 0 || f > 3)
        return 0;
    uint32_t w[8] = {1, 0, 0, 0, 0, 0, 0, 0};      // 2^p in base 2^32, little end first
    for (int i = 0; i < kPrec[f]; i++) {
        uint32_t carry = 0;
        for (int k = 0; k < 8; k++) {
            const uint64_t v = ((uint64_t)w[k] << 1) | carry;
            w[k] = (uint32_t)v;
            carry = (uint32_t)(v >> 32);
        }
    }
    size_t digits = 0;
    for (;;) {
        bool zero = true;
        uint64_t rem = 0;
        for (int k = 7; k >= 0; k--) {
            const uint64_t cur = (rem << 32) | w[k];
            w[k] = (uint32_t)(cur / 10);
            rem = cur % 10;
            zero &= w[k] == 0;
        }
        digits++;
        if (zero)
            break;
    }
    return 1 + digits;
}

#define CHARS_STATUS(who) off_path(who, "5.12 character conversion")
extern "C" CFT_API cft_status cft_from_decimal_char(cft_device *, cft_format, cft_round,
                                                    const char *const *, void *, size_t,
                                                    size_t *, uint32_t *)
{
    return CHARS_STATUS("cft_from_decimal_char");
}

extern "C" CFT_API cft_status cft_to_decimal_char(cft_device *, cft_format, cft_round,
                                                  const void *, size_t, char *, size_t, size_t *,
                                                  uint32_t *)
{
    return CHARS_STATUS("cft_to_decimal_char");
}

extern "C" CFT_API cft_status cft_from_hex_char(cft_device *, cft_format, cft_round,
                                                const char *const *, void *, size_t, size_t *,
                                                uint32_t *)
{
    return CHARS_STATUS("cft_from_hex_char");
}

extern "C" CFT_API cft_status cft_to_hex_char(cft_device *, cft_format, const void *, char *,
                                              size_t, size_t *)
{
    return CHARS_STATUS("cft_to_hex_char");
}
#undef CHARS_STATUS

#define PAYLOAD(name)                                                                       \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, const void *, void *,      \
                                       size_t)                                              \
    {                                                                                       \
        return off_path(#name, "the 9.7 NaN payload operation");                           \
    }
PAYLOAD(cft_get_payload) PAYLOAD(cft_set_payload) PAYLOAD(cft_set_payload_signaling)
#undef PAYLOAD

#define MAG(name)                                                                           \
    extern "C" CFT_API cft_status name(cft_device *, cft_format, const void *, const void *, \
                                       void *, size_t, uint32_t *)                          \
    {                                                                                       \
        return off_path(#name, "the 9.6 magnitude operation");                             \
    }
MAG(cft_min_mag) MAG(cft_max_mag) MAG(cft_minnum_mag) MAG(cft_maxnum_mag)
#undef MAG

extern "C" CFT_API int cft_is754version1985(void) { return 0; }
extern "C" CFT_API int cft_is754version2008(void) { return 0; }
extern "C" CFT_API int cft_is754version2019(void) { return 0; }
