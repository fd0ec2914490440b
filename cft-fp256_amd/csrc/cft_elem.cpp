// cft_run, cft_run_ex, cft_div and cft_sqrt: argument rules on the host, then
// one kernel launch on the device handle's stream.
//
// The argument rules and their order are the reference's run_impl
// (host/src/device.c:1699-2174), cft_run_ex (device.c:2490-2600) and
// divsqrt_validate (host/src/divsqrt.c:1574-1600): NULL operand an opcode
// reads refused, operands it does not read read as +0, d may alias a, b or c
// on a dense run, n == 0 is OK with no flags, reductions are refused here.
// What replaces the reference's chunked pinned staging (host/src/
// backend_hip.cpp:1603-1757) is operand binding: an operand in a resident
// buffer is read in place on the GPU, any other is copied once.
#include "cft_internal.hpp"
#include "kernels.h"

#include <string.h>

#include <algorithm>
#include <vector>

namespace {

bool is_reduction(int op) { return op == 24 || op == 25 || op == 28 || op == 29 || op == 31; }

unsigned op_operands(int op) { return cft_op_reads(op); }

// Everything after argument checking: bind, launch, read back.  which:
// -1 an elementwise opcode, 0 div, 1 sqrt.
cft_status execute(const char *who, cft_device *dev, int which, int op, int fmt, int rnd,
                   const void *a, const void *b, const void *c, void *d, size_t n,
                   uint32_t smask, uint32_t *flags_out)
{
    cft_status st = dev_enter(dev, who);
    if (st != CFT_OK)
        return st;
    const size_t esz = cft_format_size((cft_format)fmt);
    const void *da = nullptr, *db = nullptr, *dc = nullptr;
    void *dd = nullptr;
    cft_buffer *dbuf = nullptr;
    const size_t na = (smask & 1u) ? 1 : n, nb = (smask & 2u) ? 1 : n, nc = (smask & 4u) ? 1 : n;
    if ((st = bind_in(dev, ROLE_A, a, na * esz, &da, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_B, b, nb * esz, &db, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_C, c, nc * esz, &dc, who)) != CFT_OK ||
        (st = bind_out(dev, ROLE_D, d, n * esz, &dd, &dbuf, who)) != CFT_OK)
        return st;
    // Inside one resident buffer an operand can share device memory with d.
    // The same window (each element read before it is written) is safe; a
    // window that overlaps d's otherwise - shifted by some elements, or a
    // scalar inside it - would race across threads, so it is copied into
    // its role's arena first and the call sees the bytes from before it, as
    // it does for operands staged from host memory.
    const uint8_t *d0 = (const uint8_t *)dd, *d1 = d0 + n * esz;
    const void **din[3] = {&da, &db, &dc};
    const size_t nin[3] = {na, nb, nc};
    for (int r = 0; r < 3; r++) {
        const uint8_t *p0 = (const uint8_t *)*din[r], *p1 = p0 + nin[r] * esz;
        if (!p0 || p1 <= d0 || p0 >= d1 || (p0 == d0 && nin[r] == n))
            continue;
        DevArena &ar = dev->stage[ROLE_A + r];
        if ((st = arena_reserve(dev, ar, nin[r] * esz, who)) != CFT_OK)
            return st;
        const hipError_t ce = hipMemcpyAsync(ar.p, p0, nin[r] * esz, hipMemcpyDeviceToDevice,
                                             dev->stream);
        if (ce != hipSuccess)
            return dev_fail(dev, ce, who, "copying an operand that overlaps d");
        *din[r] = ar.p;
    }
    // The kernel ORs its flags into the handle's flag slots (zero between
    // calls); the gather launch after it ORs the slots into the pinned host
    // word and zeroes them again: no clearing launch before, no copy after.
    const int grid = grid_for(dev, n);
    const int nslots = std::min(grid, (int)CFTK_FLAG_SLOTS);
    const uint32_t slot_mask = (uint32_t)(nslots > 1 ? CFTK_FLAG_SLOTS - 1 : 0);
    hipError_t e = timer_begin(dev);
    if (e == hipSuccess)
        e = which < 0 ? cftk_elem(fmt, op, rnd, da, db, dc, dd, n, smask, dev->dslots, slot_mask,
                                  grid, dev->stream)
                      : cftk_divsqrt(fmt, which, rnd, da, db, dd, n, smask, dev->dslots,
                                     slot_mask, grid, dev->stream);
    if (e == hipSuccess)
        e = timer_end(dev);
    if (e == hipSuccess)
        e = cftk_flags_gather(dev->dslots, nslots, dev->hwords_dev, 1, dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "launching the kernel");
    if (!dbuf) {
        e = hipMemcpyAsync(d, dd, n * esz, hipMemcpyDeviceToHost, dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "copying the result back");
    }
    e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "running the kernel");
    timer_collect(dev);
    if (dbuf)
        dbuf->host_valid = false;               // the device copy is the truth now
    cft_emit_flags(dev, ((volatile uint32_t *)dev->hwords)[0], flags_out);
    return CFT_OK;
}

// host/src/device.c:1699-1800, the checks before any work.
cft_status check_common(const char *who, cft_device *dev, int op, int fmt, int rnd)
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
    if (op < 0 || op > 255) {
        cft_set_error("%s: opcode %d does not fit the device's opcode byte (0 to 255)", who,
                      op);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (is_reduction(op)) {
        cft_set_error("%s: opcode %d (%s) is a reduction - n inputs, one output - and is "
                      "issued through cft_reduce, which %s does not take",
                      who, op, cft_op_name((cft_op)op), who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

struct Tables {
    const uint32_t *idx[3];
    size_t src[3];
};

bool windows_overlap(const void *p, size_t pb, const void *q, size_t qb)
{
    const uint8_t *x = (const uint8_t *)p, *y = (const uint8_t *)q;
    return x < y + qb && y < x + pb;
}

cft_status run_impl(const char *who, cft_device *dev, int op, int fmt, int rnd, const void *a,
                    const void *b, const void *c, void *d, size_t n, uint32_t smask,
                    const Tables *tab, uint32_t *flags_out, uint32_t *bus_out)
{
    if (bus_out)
        *bus_out = 0;
    cft_status st = check_common(who, dev, op, fmt, rnd);
    if (st != CFT_OK)
        return st;
    if (n == 0) {
        cft_emit_flags(dev, 0, flags_out);
        return CFT_OK;
    }
    if (!d) {
        cft_set_error("%s: d is NULL, and %zu elements are to be written to it", who, n);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const unsigned need = op_operands(op);
    if (((need & 1u) && !a) || ((need & 2u) && !b) || ((need & 4u) && !c)) {
        cft_set_error("%s: opcode %d (%s) reads operand %c, which is NULL", who, op,
                      cft_op_name((cft_op)op),
                      ((need & 1u) && !a) ? 'a' : ((need & 2u) && !b) ? 'b' : 'c');
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const size_t esz = cft_format_size((cft_format)fmt);
    if (n > ((size_t)-1) / esz) {
        cft_set_error("%s: n = %zu elements of %zu bytes are more bytes than a size_t counts",
                      who, n, esz);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    // operands the opcode does not read are +0 (NULL) on the device
    if (!(need & 1u)) a = nullptr;
    if (!(need & 2u)) b = nullptr;
    if (!(need & 4u)) c = nullptr;

    const bool tables = tab && (tab->idx[0] || tab->idx[1] || tab->idx[2]);
    if (!tables)
        return execute(who, dev, -1, op, fmt, rnd, a, b, c, d, n, smask, flags_out);

    // Index tables (ABI 0.14): d may overlap no operand window; every index
    // below its source's length or CFT_IDX_NONE; then the operand is
    // gathered and the run is the dense one (host/src/device.c:1836-1893).
    const void *opnd[3] = {a, b, c};
    for (int r = 0; r < 3; r++) {
        if (!opnd[r])
            continue;
        const size_t obytes =
            (tab->idx[r] ? tab->src[r] : ((smask >> r) & 1u) ? 1u : n) * esz;
        if (windows_overlap(d, n * esz, opnd[r], obytes)) {
            cft_set_error("%s: d overlaps operand %c, and a run with an index table may not "
                          "write over an operand it reads", who, 'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    for (int r = 0; r < 3; r++) {
        if (!tab->idx[r])
            continue;
        for (size_t e = 0; e < n; e++) {
            const uint32_t ix = tab->idx[r][e];
            if (ix != CFT_IDX_NONE && (size_t)ix >= tab->src[r]) {
                cft_set_error("%s: idx_%c[%zu] = %u is past its source's %zu elements "
                              "(idx_%c_src)", who, 'a' + r, e, ix, tab->src[r], 'a' + r);
                return CFT_ERR_INVALID_ARGUMENT;
            }
        }
    }
    std::vector<uint8_t> g[3];
    const void *dense[3] = {a, b, c};
    for (int r = 0; r < 3; r++) {
        if (!tab->idx[r] || !opnd[r])
            continue;
        g[r].assign(n * esz, 0);
        const uint8_t *src = (const uint8_t *)opnd[r];
        if (cft_buffer *bf = find_buffer(dev, src, tab->src[r] * esz)) {
            if ((st = cft_buffer_from_device(bf)) != CFT_OK)
                return st;
        }
        for (size_t e = 0; e < n; e++) {
            const uint32_t ix = tab->idx[r][e];
            if (ix != CFT_IDX_NONE)
                memcpy(&g[r][e * esz], src + (size_t)ix * esz, esz);
        }
        dense[r] = g[r].data();
    }
    return execute(who, dev, -1, op, fmt, rnd, dense[0], dense[1], dense[2], d, n, smask,
                   flags_out);
}

} // namespace

extern "C" CFT_API cft_status cft_run(cft_device *dev, cft_op op, cft_format fmt, cft_round rnd,
                                      const void *a, const void *b, const void *c, void *d,
                                      size_t n, uint32_t *flags_out, uint32_t *bus_out)
{
    cft_clear_error();
    return run_impl("cft_run", dev, (int)op, (int)fmt, (int)rnd, a, b, c, d, n, 0u, nullptr,
                    flags_out, bus_out);
}

extern "C" CFT_API cft_status cft_run_ex(cft_device *dev, cft_op op, cft_format fmt,
                                         cft_round rnd, const cft_elem_args *args)
{
    cft_clear_error();
    if (!dev || !args) {
        cft_set_error("cft_run_ex: %s is NULL", dev ? "args" : "dev");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (args->struct_size != sizeof(cft_elem_args)) {
        cft_set_error("cft_run_ex was given a %zu-byte cft_elem_args and this library's is %zu "
                      "bytes; an input struct of another size is refused",
                      args->struct_size, sizeof(cft_elem_args));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (args->scalar_mask & ~7u) {
        cft_set_error("cft_run_ex: scalar_mask 0x%x sets bits above the three operands",
                      args->scalar_mask);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (((args->scalar_mask & 1u) && !args->a) || ((args->scalar_mask & 2u) && !args->b) ||
        ((args->scalar_mask & 4u) && !args->c)) {
        cft_set_error("cft_run_ex: scalar_mask names operand %c as one element and it is NULL",
                      ((args->scalar_mask & 1u) && !args->a) ? 'a'
                      : ((args->scalar_mask & 2u) && !args->b) ? 'b' : 'c');
        return CFT_ERR_INVALID_ARGUMENT;
    }
    Tables tab;
    tab.idx[0] = args->idx_a; tab.idx[1] = args->idx_b; tab.idx[2] = args->idx_c;
    tab.src[0] = args->idx_a_src; tab.src[1] = args->idx_b_src; tab.src[2] = args->idx_c_src;
    const void *opnd[3] = {args->a, args->b, args->c};
    for (int r = 0; r < 3; r++) {
        if (!tab.idx[r]) {
            if (tab.src[r]) {
                cft_set_error("cft_run_ex: idx_%c_src = %zu names a source length and there "
                              "is no idx_%c", 'a' + r, tab.src[r], 'a' + r);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            continue;
        }
        if (!opnd[r]) {
            cft_set_error("cft_run_ex: idx_%c indexes operand %c, which is NULL", 'a' + r,
                          'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (args->scalar_mask & (1u << r)) {
            cft_set_error("cft_run_ex: operand %c is both scalar and indexed", 'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (tab.src[r] == 0) {
            cft_set_error("cft_run_ex: idx_%c is set and idx_%c_src is 0", 'a' + r, 'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    if ((int)op >= 0 && (int)op <= 255 && !is_reduction((int)op)) {
        const unsigned need = op_operands((int)op);
        for (int r = 0; r < 3; r++)
            if (tab.idx[r] && !((need >> r) & 1u)) {
                cft_set_error("cft_run_ex: idx_%c indexes operand %c, which opcode %d (%s) does "
                              "not read", 'a' + r, 'a' + r, (int)op, cft_op_name(op));
                return CFT_ERR_INVALID_ARGUMENT;
            }
    }
    return run_impl("cft_run_ex", dev, (int)op, (int)fmt, (int)rnd, args->a, args->b, args->c,
                    args->d, args->n, args->scalar_mask, &tab, args->flags_out, args->bus_out);
}

// host/src/divsqrt.c:1574-1600
static cft_status divsqrt_impl(const char *who, cft_device *dev, int fmt, int rnd, int which,
                               const void *a, const void *b, void *d, size_t n,
                               uint32_t *flags_out, uint32_t *bus_out)
{
    if (bus_out)
        *bus_out = 0;
    if (!dev) {
        cft_set_error("%s: dev is NULL", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (fmt < 0 || fmt > 3) {
        cft_set_error("%s: format %d is not a cft_format", who, fmt);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (rnd < 0 || rnd > 4) {
        cft_set_error("%s: rounding attribute %d is not a cft_round", who, rnd);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (n == 0) {
        cft_emit_flags(dev, 0, flags_out);
        return CFT_OK;
    }
    if (!a || !d || (which == 0 && !b)) {
        cft_set_error("%s: %s is NULL", who, !a ? "a" : !d ? "d" : "b");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const size_t esz = cft_format_size((cft_format)fmt);
    if (n > ((size_t)-1) / esz) {
        cft_set_error("%s: n = %zu elements overflow a size_t", who, n);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return execute(who, dev, which, 0, fmt, rnd, a, which == 0 ? b : nullptr, nullptr, d, n, 0u,
                   flags_out);
}

extern "C" CFT_API cft_status cft_div(cft_device *dev, cft_format fmt, cft_round rnd,
                                      const void *a, const void *b, void *d, size_t n,
                                      uint32_t *flags_out, uint32_t *bus_out)
{
    cft_clear_error();
    return divsqrt_impl("cft_div", dev, (int)fmt, (int)rnd, 0, a, b, d, n, flags_out, bus_out);
}

extern "C" CFT_API cft_status cft_sqrt(cft_device *dev, cft_format fmt, cft_round rnd,
                                       const void *a, void *d, size_t n, uint32_t *flags_out,
                                       uint32_t *bus_out)
{
    cft_clear_error();
    return divsqrt_impl("cft_sqrt", dev, (int)fmt, (int)rnd, 1, a, nullptr, d, n, flags_out,
                        bus_out);
}
