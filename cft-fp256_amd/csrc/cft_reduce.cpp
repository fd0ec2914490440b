// cft_reduce and cft_reduce_seg: argument rules on the host, then the tree on
// the device, a few levels a launch (k_reduce.hip), every segment at once.
//
// The argument rules and their order are the reference's reduce_body
// (host/src/device.c:2759-2885) and reduce_seg_body (device.c:3251-3345): dev,
// format, rounding attribute, a reduction opcode, d, then for the segmented
// form seg >= 1 and a whole number of segments; n == 0 is the identity (+0,
// -inf for CFT_MAXALL) for cft_reduce and nothing for cft_reduce_seg, both
// raising nothing; an operand the opcode reads must not be NULL.  What is
// computed is the contract of host/include/cft.h:1210-1340: the fixed tree
// whose left child is the largest power of two strictly inside the range
// (host/src/softfloat.c:880-925), leaves a[i], round(a[i]*b[i]),
// round(a[i]*a[i]) or |a[i]|, 754-2019 maximum for CFT_MAXALL, and for
// sumSquare / sumAbs 9.4's infinity-ahead-of-NaN row (device.c:2678-2714),
// applied here per segment from what the first launch recorded.
#include "cft_internal.hpp"
#include "kernels.h"

#include <string.h>

#include <vector>

namespace {

enum { OP_SUM = 24, OP_DOT = 25, OP_SUMSQ = 28, OP_SUMABS = 29, OP_MAXALL = 31 };

bool is_reduction(int op)
{
    return op == OP_SUM || op == OP_DOT || op == OP_SUMSQ || op == OP_SUMABS ||
           op == OP_MAXALL;
}

// The encoding of +inf (sign 0) or -inf (sign 1) in format fmt.
void store_inf(int fmt, int sign, uint8_t *out)
{
    static const int kExpBits[4] = {8, 11, 15, 19};
    const int esz = (int)cft_format_size((cft_format)fmt);
    const int width = esz * 8, man = width - 1 - kExpBits[fmt];
    memset(out, 0, (size_t)esz);
    for (int i = man; i < width - 1; i++)
        out[i / 8] |= (uint8_t)(1u << (i % 8));
    if (sign)
        out[esz - 1] |= 0x80u;
}

bool is_nan(int fmt, const uint8_t *p)
{
    static const int kExpBits[4] = {8, 11, 15, 19};
    const int esz = (int)cft_format_size((cft_format)fmt);
    const int width = esz * 8, man = width - 1 - kExpBits[fmt];
    bool frac = false;
    for (int i = man; i < width - 1; i++)
        if (!((p[i / 8] >> (i % 8)) & 1))
            return false;
    for (int i = 0; i < man && !frac; i++)
        frac = (p[i / 8] >> (i % 8)) & 1;
    return frac;
}

// host/src/device.c:2759-2822 and 3251-3318: the checks every call makes
// before anything is read.
cft_status check_head(const char *who, cft_device *dev, int op, int fmt, int rnd, void *d,
                      bool seg_form)
{
    if (!dev) {
        cft_set_error("%s: dev is NULL", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (fmt < 0 || fmt > 3) {
        cft_set_error("%s: format %d is not a cft_format (CFT_FP32 to CFT_FP256, 0 to 3)",
                      who, fmt);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (rnd < 0 || rnd > 4) {
        cft_set_error("%s: rounding attribute %d is not a cft_round (CFT_RNE to CFT_RMM, 0 "
                      "to 4)", who, rnd);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!is_reduction(op)) {
        cft_set_error("%s: opcode %d (%s) is not a reduction, and elementwise opcodes are "
                      "issued through cft_run", who, op, cft_op_name((cft_op)op));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!d) {
        cft_set_error(seg_form ? "%s: d is NULL, and one result per segment is written there"
                               : "%s: d is NULL, and the result is one element written there",
                      who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

// host/src/device.c:2849-2862: an operand the opcode reads, and a byte count
// a size_t holds.
cft_status check_operands(const char *who, int op, const void *a, const void *b, size_t n,
                          size_t esz)
{
    const bool need_b = op == OP_DOT;
    if (!a || (need_b && !b)) {
        cft_set_error("%s: opcode %d (%s) reads operand %c, which is NULL", who, op,
                      cft_op_name((cft_op)op), !a ? 'a' : 'b');
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (n > ((size_t)-1) / esz) {
        cft_set_error("%s: n = %zu elements of %zu bytes are more bytes than a size_t counts",
                      who, n, esz);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

// Write `bytes` of host data over d, which may lie in a resident buffer: the
// copy (or copies) holding the current bytes get them, so the buffer's
// authority rule is unchanged.
cft_status put_host(cft_device *dev, void *d, const void *src, size_t bytes, const char *who)
{
    cft_buffer *buf = find_buffer(dev, d, bytes);
    if (!buf) {
        memmove(d, src, bytes);
        return CFT_OK;
    }
    if (buf->dev_valid) {
        cft_status st = dev_enter(dev, who);
        if (st != CFT_OK)
            return st;
        void *dp = (uint8_t *)buf->devp + ((uint8_t *)d - (uint8_t *)buf->host);
        hipError_t e = hipMemcpyAsync(dp, src, bytes, hipMemcpyHostToDevice, dev->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "writing the result to the buffer's device copy");
    }
    if (buf->host_valid)
        memmove(d, src, bytes);
    return CFT_OK;
}

// The reduction over nseg segments of seg elements, n = nseg * seg >= 1.
cft_status execute(const char *who, cft_device *dev, int op, int fmt, int rnd, const void *a,
                   const void *b, void *d, size_t nseg, size_t seg, uint32_t *flags_out)
{
    cft_status st = dev_enter(dev, who);
    if (st != CFT_OK)
        return st;
    const size_t esz = cft_format_size((cft_format)fmt), n = nseg * seg;
    const bool override94 = op == OP_SUMSQ || op == OP_SUMABS;
    const void *da = nullptr, *db = nullptr;
    if ((st = bind_in(dev, ROLE_A, a, n * esz, &da, who)) != CFT_OK)
        return st;
    if (op == OP_DOT && (st = bind_in(dev, ROLE_B, b, n * esz, &db, who)) != CFT_OK)
        return st;
    // the partials: the first launch writes nseg * ceil(seg/16), each later
    // one a sixteenth of the one before, ping-ponging between two arenas
    const size_t first = nseg * ((seg + 15) / 16);
    const size_t second = nseg * ((((seg + 15) / 16) + 15) / 16);
    if ((st = arena_reserve(dev, dev->stage[ROLE_RED0], first * esz, who)) != CFT_OK ||
        (st = arena_reserve(dev, dev->stage[ROLE_RED1], second * esz, who)) != CFT_OK)
        return st;
    // SUMSQ/SUMABS: a flag word and a specials word a segment, cleared here
    // and read back.  The others: the call's flags through the handle's flag
    // slots (zero between calls), gathered into the pinned host word.
    const size_t nwords = override94 ? 2 * nseg : 0;
    uint32_t *segflags = dev->dslots, *segspec = nullptr;
    if (override94) {
        if ((st = arena_reserve(dev, dev->stage[ROLE_RWORDS], nwords * 4, who)) != CFT_OK)
            return st;
        segflags = (uint32_t *)dev->stage[ROLE_RWORDS].p;
        segspec = segflags + nseg;
    }
    void *part[2] = {dev->stage[ROLE_RED0].p, dev->stage[ROLE_RED1].p};

    hipError_t e = override94 ? cftk_fill32(segflags, 0u, nwords, dev->stream) : hipSuccess;
    if (e == hipSuccess)
        e = timer_begin(dev);
    const int leaf = op == OP_DOT ? 1 : op == OP_SUMABS ? 2 : op == OP_SUMSQ ? 3 : 0;
    const int kind = op == OP_MAXALL ? 1 : 0;
    size_t m = seg;
    const void *src = da;
    int out = 0;
    for (int pass = 0; e == hipSuccess; pass++) {
        const size_t m2 = (m + 15) / 16;
        e = cftk_reduce(fmt, src, pass == 0 ? db : nullptr, part[out], nseg, m, m2,
                        pass == 0 ? leaf : 0, kind, rnd, override94 ? 1 : 0, segflags,
                        pass == 0 ? segspec : nullptr, grid_for(dev, nseg * m2), dev->stream);
        src = part[out];
        if (m2 == 1)
            break;
        out ^= 1;
        m = m2;
    }
    if (e == hipSuccess)
        e = timer_end(dev);
    if (e == hipSuccess && !override94)
        e = cftk_flags_gather(dev->dslots, CFTK_FLAG_SLOTS, dev->hwords_dev, 1, dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "launching the reduction");

    // results home (or into d's device copy), then the flag words
    cft_buffer *dbuf = find_buffer(dev, d, nseg * esz);
    std::vector<uint8_t> res;
    void *dd = nullptr;
    if (dbuf) {
        cft_buffer *bb = nullptr;
        if ((st = bind_out(dev, ROLE_D, d, nseg * esz, &dd, &bb, who)) != CFT_OK)
            return st;
        e = hipMemcpyAsync(dd, src, nseg * esz, hipMemcpyDeviceToDevice, dev->stream);
    } else if (!override94) {
        e = hipMemcpyAsync(d, src, nseg * esz, hipMemcpyDeviceToHost, dev->stream);
    }
    if (e == hipSuccess && override94) {
        res.resize(nseg * esz);
        e = hipMemcpyAsync(res.data(), src, nseg * esz, hipMemcpyDeviceToHost, dev->stream);
    }
    std::vector<uint32_t> words(nwords);
    if (e == hipSuccess && override94)
        e = hipMemcpyAsync(words.data(), segflags, nwords * 4, hipMemcpyDeviceToHost,
                           dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "running the reduction");
    timer_collect(dev);
    if (dbuf)
        dbuf->host_valid = false;                   // the device copy is the truth now

    uint32_t acc = 0;
    if (!override94) {
        acc = ((volatile uint32_t *)dev->hwords)[0];
    } else {
        // 754-2019 9.4: a NaN result over a segment holding an infinity is
        // +inf, with INVALID exactly when the segment holds a signalling NaN
        bool changed = false;
        for (size_t s = 0; s < nseg; s++) {
            uint32_t fl = words[s];
            const uint32_t spec = words[nseg + s];
            if ((spec & 1u) && is_nan(fmt, res.data() + s * esz)) {
                store_inf(fmt, 0, res.data() + s * esz);
                fl = (spec & 2u) ? CFT_FLAG_INVALID : 0u;
                changed = true;
            }
            acc |= fl;
        }
        if (!dbuf) {
            memcpy(d, res.data(), nseg * esz);
        } else if (changed) {
            e = hipMemcpyAsync(dd, res.data(), nseg * esz, hipMemcpyHostToDevice, dev->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(dev->stream);
            if (e != hipSuccess)
                return dev_fail(dev, e, who, "writing the result to the buffer's device copy");
        }
    }
    cft_emit_flags(dev, acc, flags_out);
    return CFT_OK;
}

} // namespace

// cft_reduce's argument rules in its order, for its variants.
cft_status reduce_check(const char *who, cft_device *dev, int op, int fmt, int rnd,
                        const void *a, const void *b, void *d, size_t n)
{
    cft_status st = check_head(who, dev, op, fmt, rnd, d, false);
    if (st == CFT_OK && n)
        st = check_operands(who, op, a, b, n, cft_format_size((cft_format)fmt));
    return st;
}

// reference host/include/cft.h:1284-1293, host/src/device.c:2731-2757
extern "C" CFT_API cft_status cft_reduce(cft_device *dev, cft_op op, cft_format fmt,
                                         cft_round rnd, const void *a, const void *b, void *d,
                                         size_t n, uint32_t *flags_out, uint32_t *bus_out)
{
    static const char *who = "cft_reduce";
    cft_clear_error();
    if (bus_out)
        *bus_out = 0;
    cft_status st = check_head(who, dev, (int)op, (int)fmt, (int)rnd, d, false);
    if (st != CFT_OK)
        return st;
    const size_t esz = cft_format_size(fmt);
    if (n == 0) {
        uint8_t z[32] = {0};
        if (op == CFT_MAXALL)
            store_inf((int)fmt, 1, z);
        if ((st = put_host(dev, d, z, esz, who)) != CFT_OK)
            return st;
        cft_emit_flags(dev, 0, flags_out);
        return CFT_OK;
    }
    if ((st = check_operands(who, (int)op, a, b, n, esz)) != CFT_OK)
        return st;
    return execute(who, dev, (int)op, (int)fmt, (int)rnd, a, b, d, 1, n, flags_out);
}

// reference host/include/cft.h:1329-1340, host/src/device.c:3219-3400
extern "C" CFT_API cft_status cft_reduce_seg(cft_device *dev, cft_op op, cft_format fmt,
                                             cft_round rnd, const void *a, const void *b,
                                             void *d, size_t n, size_t seg,
                                             uint32_t *flags_out, uint32_t *bus_out)
{
    static const char *who = "cft_reduce_seg";
    cft_clear_error();
    if (bus_out)
        *bus_out = 0;
    cft_status st = check_head(who, dev, (int)op, (int)fmt, (int)rnd, d, true);
    if (st != CFT_OK)
        return st;
    if (seg == 0) {
        cft_set_error("%s: a segment is at least one element", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (n % seg) {
        cft_set_error("%s: n = %zu is not a whole number of segments of %zu - the last %zu "
                      "elements would belong to no result",
                      who, n, seg, n % seg);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (n == 0) {
        cft_emit_flags(dev, 0, flags_out);
        return CFT_OK;
    }
    const size_t esz = cft_format_size(fmt);
    if ((st = check_operands(who, (int)op, a, b, n, esz)) != CFT_OK)
        return st;
    return execute(who, dev, (int)op, (int)fmt, (int)rnd, a, b, d, n / seg, seg, flags_out);
}
