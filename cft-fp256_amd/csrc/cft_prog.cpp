// Programs: cft_program_load (parse, validate, pre-decode), the runs
// (cft_program_run / _run_bank / _run_ex), cft_program_get_info,
// cft_program_digest and cft_program_free.
//
// The loader's refusals, their statuses and their order are the reference's
// (host/src/program.c): the header (program.c:1390-1470), the device's caps
// for deposits, constants and the scratch block (program.c:1480-1530),
// seq_validate (program.c:601-803, with seq_check_operands at 201-310 and
// the control-code field table at 464-540), the header against the device
// (program.c:1160-1280) and the stream against the device (program.c:
// 959-1100).  The run's argument rules are cft_program_run_ex and the checks
// it calls (program.c:1640-1770, 2050-2250, 2258-2330) and the wrappers at
// program.c:2340-2420.
//
// What is this build's own: after validation the image is decoded once into
// 16-byte records (seq_code.h) with the registers renamed densely, and the
// constant bank of an image that carries one is uploaded once, so a run
// only binds its streams and launches the interpreter (k_seq.hip).
//
// This device does not publish CFT_SEQ_FEAT_AUGADD: an image that uses
// control codes 10 or 11 is refused at load by name, as the reference
// refuses it on any device without the bit (program.c:377-388).
#include "seq_prog.hpp"

#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr uint32_t kMagic = 0x50544643u;          // "CFTP"
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 32;
constexpr uint32_t kFlagsKnown =
    CFT_PROG_FLAG_BANK_EXT | CFT_PROG_FLAG_SCRATCH_IO | CFT_PROG_FLAG_SCRATCH_STRICT;
constexpr uint64_t kMaxIssue = 1ull << 40;       // worst-case instruction count
constexpr uint64_t kCountCap = 1ull << 62;
constexpr int kMaxLoops = 4, kMaxQuiet = 4;
constexpr uint32_t kImmReserved = 0x80000000u, kKxIndices = 0x00FFFFFFu,
                   kKx9Mask = 0x70000000u, kStepMask = 0x00000FFFu;

uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

// One instruction word, its fields as the encoding holds them
// (host/src/seq_exec.h:175-196): the four-bit register fields and their
// fifth bits kept apart, because a constant operand's fifth bit is a
// reserved bit rather than part of an index.
struct Dec {
    int op, rd, ra, rb, rc, rnd;
    int hd, ha, hb, hc;
    int k9a, k9b, k9c;
    int ka, kb, kc, kx, ctrl;
    uint32_t imm;
};

Dec decode(uint64_t w)
{
    Dec d;
    d.op = (int)(w & 0xFF);
    d.rd = (int)((w >> 8) & 0xF);
    d.ra = (int)((w >> 12) & 0xF);
    d.rb = (int)((w >> 16) & 0xF);
    d.rc = (int)((w >> 20) & 0xF);
    d.rnd = (int)((w >> 24) & 7);
    d.ka = (int)((w >> 27) & 1);
    d.kb = (int)((w >> 28) & 1);
    d.kc = (int)((w >> 29) & 1);
    d.kx = (int)((w >> 30) & 1);
    d.ctrl = (int)((w >> 31) & 1);
    d.imm = (uint32_t)(w >> 32);
    d.hd = (int)((d.imm >> 24) & 1);
    d.ha = (int)((d.imm >> 25) & 1);
    d.hb = (int)((d.imm >> 26) & 1);
    d.hc = (int)((d.imm >> 27) & 1);
    d.k9a = (int)((d.imm >> 28) & 1);
    d.k9b = (int)((d.imm >> 29) & 1);
    d.k9c = (int)((d.imm >> 30) & 1);
    return d;
}

int reg5(int lo, int hi) { return lo | (hi << 4); }

// Constant index of operand k (0 a, 1 b, 2 c) under the instruction's
// encoding: the field itself, or under kx a byte of imm and its ninth bit.
uint32_t konst_index(const Dec &d, int k)
{
    const int field = k == 0 ? d.ra : k == 1 ? d.rb : d.rc;
    const int k9 = k == 0 ? d.k9a : k == 1 ? d.k9b : d.k9c;
    if (!d.kx)
        return (uint32_t)field;
    return ((d.imm >> (8 * k)) & 0xFFu) | ((uint32_t)k9 << 8);
}

// host/src/device.c:1287-1301: a program past one of the device's published
// capacities.
cft_status cap_refusal(const char *field, unsigned long asked, unsigned long cap,
                       const char *units, const char *caps_field)
{
    cft_set_error("this program's %s is %lu and this device's is %lu (%s); cft_get_caps "
                  "publishes it as cft_caps.%s", field, asked, cap, units, caps_field);
    return CFT_ERR_UNSUPPORTED;
}

const char *ctrl_name(int op)
{
    static const char *const names[] = {"HALT", "REPEAT", "ENDREP", "DEPOSIT", "SETACT",
                                        "ACTALL", "STL", "LDL", "STX", "LDX", "AUGADD",
                                        "AUGERR", "QUIET", "ENDQUIET", "RAISE"};
    return op >= 0 && op < 15 ? names[op] : "?";
}

// What each control code reads (program.c:464-486): the operand fields
// (bit 0 rd, 1 ra, 2 rb, 3 rc) and the bits of imm.
struct CtrlRow {
    unsigned regs;
    uint32_t imm;
};
constexpr uint32_t HI_RD = 1u << 24, HI_RA = 1u << 25, HI_RB = 1u << 26;
constexpr CtrlRow kCtrl[15] = {
    {0, 0},                                          // HALT
    {0, 0xFFFFFFFFu},                                // REPEAT
    {0, 0},                                          // ENDREP
    {2, HI_RA},                                      // DEPOSIT
    {2, HI_RA},                                      // SETACT
    {0, 0},                                          // ACTALL
    {2, kKxIndices | HI_RA},                         // STL
    {1, kKxIndices | HI_RD},                         // LDL
    {2 | 4, HI_RA | HI_RB | kStepMask},              // STX
    {1 | 4, HI_RD | HI_RB | kStepMask},              // LDX
    {1 | 2 | 4, HI_RD | HI_RA | HI_RB},              // AUGADD
    {1 | 2 | 4, HI_RD | HI_RA | HI_RB},              // AUGERR
    {0, 0},                                          // QUIET
    {0, 0},                                          // ENDQUIET
    {2, HI_RA},                                      // RAISE
};

// ---- SHA-256 (FIPS 180-4), for cft_program_digest ------------------------

struct Sha256 {
    uint32_t h[8];
    uint8_t buf[64];
    uint64_t len = 0;
    size_t fill = 0;

    Sha256()
    {
        static const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                       0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
        memcpy(h, iv, sizeof h);
    }
    static uint32_t ror(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }
    void block(const uint8_t *p)
    {
        static const uint32_t K[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1,
            0x923f82a4, 0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3,
            0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786,
            0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
            0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147,
            0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13,
            0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
            0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
            0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a,
            0x5b9cca4f, 0x682e6ff3, 0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208,
            0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
        uint32_t w[64];
        for (int i = 0; i < 16; i++)
            w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
                   ((uint32_t)p[4 * i + 2] << 8) | p[4 * i + 3];
        for (int i = 16; i < 64; i++) {
            const uint32_t s0 = ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6],
                 hh = h[7];
        for (int i = 0; i < 64; i++) {
            const uint32_t t1 = hh + (ror(e, 6) ^ ror(e, 11) ^ ror(e, 25)) +
                                ((e & f) ^ (~e & g)) + K[i] + w[i];
            const uint32_t t2 = (ror(a, 2) ^ ror(a, 13) ^ ror(a, 22)) +
                                ((a & b) ^ (a & c) ^ (b & c));
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d;
        h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
    void update(const void *data, size_t n)
    {
        const uint8_t *p = (const uint8_t *)data;
        len += n;
        while (n) {
            const size_t k = std::min(n, 64 - fill);
            memcpy(buf + fill, p, k);
            fill += k;
            p += k;
            n -= k;
            if (fill == 64) {
                block(buf);
                fill = 0;
            }
        }
    }
    void final(uint8_t out[32])
    {
        const uint64_t bits = len * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56)
            update(&zero, 1);
        uint8_t lb[8];
        for (int i = 0; i < 8; i++)
            lb[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(lb, 8);
        for (int i = 0; i < 8; i++) {
            out[4 * i] = (uint8_t)(h[i] >> 24);
            out[4 * i + 1] = (uint8_t)(h[i] >> 16);
            out[4 * i + 2] = (uint8_t)(h[i] >> 8);
            out[4 * i + 3] = (uint8_t)h[i];
        }
    }
};

} // namespace

namespace {

// seq_check_operands (program.c:201-310), for an ALU instruction.
cft_status check_operands(const cft_program *p, const Dec &d, uint32_t pc)
{
    static const char *const nm[3] = {"ra", "rb", "rc"};
    const int reg[3] = {d.ra, d.rb, d.rc};
    const int hi[3] = {d.ha, d.hb, d.hc};
    const int k9[3] = {d.k9a, d.k9b, d.k9c};
    const int kf[3] = {d.ka, d.kb, d.kc};
    const unsigned long at = pc;
    const char *name = cft_op_name((cft_op)d.op);

    if (d.kx && !(d.ka || d.kb || d.kc)) {
        cft_set_error("instruction %lu (%s, opcode %d) sets kx with no constant operand; "
                      "the bit would select nothing", at, name, d.op);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (d.imm & kImmReserved) {
        cft_set_error("instruction %lu (%s, opcode %d): imm[31] is reserved and set", at,
                      name, d.op);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!d.kx && (d.imm & kKxIndices)) {
        cft_set_error("instruction %lu (%s, opcode %d) has no kx, so imm[23:0] must be zero; "
                      "it is 0x%06lx", at, name, d.op,
                      (unsigned long)(d.imm & kKxIndices));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!d.kx && (d.imm & kKx9Mask)) {
        cft_set_error("instruction %lu (%s, opcode %d) has no kx, so imm[30:28] must be "
                      "zero", at, name, d.op);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    for (int k = 0; k < 3; k++)
        if (kf[k] && hi[k]) {
            cft_set_error("instruction %lu (%s, opcode %d): %s is a constant, so its "
                          "register high bit imm[%d] must be zero", at, name, d.op, nm[k],
                          25 + k);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    for (int k = 0; k < 3; k++) {
        const uint32_t byte = (d.imm >> (8 * k)) & 0xFFu;
        uint32_t idx;
        if (d.kx && kf[k]) {
            idx = byte | ((uint32_t)k9[k] << 8);
            if (reg[k]) {
                cft_set_error("instruction %lu (%s, opcode %d): %s names constant %lu "
                              "through imm, so its field must be zero; it is %d", at, name,
                              d.op, nm[k], (unsigned long)idx, reg[k]);
                return CFT_ERR_INVALID_ARGUMENT;
            }
        } else if (d.kx) {
            if (byte) {
                cft_set_error("instruction %lu (%s, opcode %d): %s is a register, so "
                              "imm[%d:%d] must be zero; it is 0x%02lx", at, name, d.op,
                              nm[k], 8 * k + 7, 8 * k, (unsigned long)byte);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            if (k9[k]) {
                cft_set_error("instruction %lu (%s, opcode %d): %s is a register, so its "
                              "ninth index bit imm[%d] must be zero", at, name, d.op, nm[k],
                              28 + k);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            continue;
        } else if (kf[k]) {
            idx = (uint32_t)reg[k];
        } else {
            continue;
        }
        if (idx >= p->n_consts) {
            cft_set_error("instruction %lu (%s, opcode %d): %s names constant %lu and the "
                          "header declares %lu", at, name, d.op, nm[k], (unsigned long)idx,
                          (unsigned long)p->n_consts);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    return CFT_OK;
}

// seq_ctrl_fields (program.c:488-540)
cft_status check_ctrl_fields(const Dec &d, uint32_t pc)
{
    static const char *const rn[4] = {"rd", "ra", "rb", "rc"};
    static const char *const fn[4] = {"ka", "kb", "kc", "kx"};
    const int regs[4] = {d.rd, d.ra, d.rb, d.rc};
    const int fl[4] = {d.ka, d.kb, d.kc, d.kx};
    const CtrlRow &row = kCtrl[d.op];
    for (int f = 0; f < 4; f++)
        if (regs[f] && !(row.regs & (1u << f))) {
            cft_set_error("instruction %lu is %s, which does not read %s; it must be zero "
                          "and is %d", (unsigned long)pc, ctrl_name(d.op), rn[f], regs[f]);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    if (d.rnd) {
        cft_set_error("instruction %lu is %s, which reads no rounding attribute; rnd must "
                      "be zero and is %d", (unsigned long)pc, ctrl_name(d.op), d.rnd);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    for (int f = 0; f < 4; f++)
        if (fl[f]) {
            cft_set_error("instruction %lu is %s, which reads no constant; %s must be "
                          "clear", (unsigned long)pc, ctrl_name(d.op), fn[f]);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    if (d.imm & ~row.imm) {
        cft_set_error("instruction %lu is %s, which reads only imm & 0x%08lx; imm is "
                      "0x%08lx", (unsigned long)pc, ctrl_name(d.op), (unsigned long)row.imm,
                      (unsigned long)d.imm);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

uint64_t count_add(uint64_t a, uint64_t b) { return std::min(a + b, kCountCap); }
uint64_t count_mul(uint64_t a, uint64_t b)
{
    return (a && b > kCountCap / a) ? kCountCap : a * b;
}

cft_status worst_refusal(uint64_t worst, uint32_t pc, bool at_end)
{
    cft_set_error("%s instruction %lu the program's worst-case instruction count is %s%llu, "
                  "past the bound of 2^40", at_end ? "ending after" : "at",
                  (unsigned long)pc, worst >= kCountCap ? "at least " : "",
                  (unsigned long long)worst);
    return CFT_ERR_INVALID_ARGUMENT;
}

// seq_validate (program.c:601-803)
cft_status validate(const cft_program *p)
{
    int depth = 0, top = 0, nb = 0, qdepth = 0;
    uint64_t mult[kMaxLoops + 1];
    uint64_t worst = 0;
    int bkind[kMaxLoops + kMaxQuiet];
    uint32_t bpc[kMaxLoops + kMaxQuiet];
    const uint32_t max_scratch = p->depth;
    mult[0] = 1;
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const Dec d = decode(p->insns[pc]);
        worst = count_add(worst, mult[top]);
        if (!d.ctrl) {
            cft_status st = check_operands(p, d, pc);
            if (st != CFT_OK)
                return st;
            if (d.rnd > 4) {
                cft_set_error("instruction %lu (%s, opcode %d): rounding attribute %d is "
                              "reserved (0 to 4 are defined)", (unsigned long)pc,
                              cft_op_name((cft_op)d.op), d.op, d.rnd);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            continue;
        }
        if (d.op >= 15) {
            cft_set_error("instruction %lu is control code %d; the control codes are 0 to 14",
                          (unsigned long)pc, d.op);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        cft_status st = check_ctrl_fields(d, pc);
        if (st != CFT_OK)
            return st;
        if ((d.op == SEQ_STL || d.op == SEQ_LDL) && max_scratch &&
            (d.imm & kKxIndices) >= max_scratch)
            return cap_refusal("scratch slot", (unsigned long)(d.imm & kKxIndices),
                               (unsigned long)max_scratch - 1u,
                               "scratch slots a lane, indexed from 0", "max_scratch");
        if (d.op == SEQ_REPEAT) {
            if (d.imm == 0) {
                cft_set_error("instruction %lu is REPEAT 0, which is not a loop",
                              (unsigned long)pc);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            if (depth >= kMaxLoops) {
                cft_set_error("instruction %lu is a REPEAT inside %d open loops; loops nest "
                              "at most %d deep", (unsigned long)pc, depth, kMaxLoops);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            depth++;
            top++;
            mult[top] = count_mul(mult[top - 1], d.imm);
            bkind[nb] = SEQ_REPEAT;
            bpc[nb++] = pc;
        } else if (d.op == SEQ_ENDREP) {
            if (depth == 0) {
                cft_set_error("instruction %lu is an ENDREP with no REPEAT open",
                              (unsigned long)pc);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            if (bkind[nb - 1] != SEQ_REPEAT) {
                cft_set_error("instruction %lu closes a loop inside the quiet region opened "
                              "at instruction %lu", (unsigned long)pc,
                              (unsigned long)bpc[nb - 1]);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            nb--;
            depth--;
            top--;
        } else if (d.op == SEQ_QUIET) {
            if (qdepth >= kMaxQuiet) {
                cft_set_error("instruction %lu is a QUIET inside %d open quiet regions; "
                              "regions nest at most %d deep", (unsigned long)pc, qdepth,
                              kMaxQuiet);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            qdepth++;
            bkind[nb] = SEQ_QUIET;
            bpc[nb++] = pc;
        } else if (d.op == SEQ_ENDQUIET) {
            if (qdepth == 0) {
                cft_set_error("instruction %lu is an ENDQUIET with no quiet region open",
                              (unsigned long)pc);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            if (bkind[nb - 1] != SEQ_QUIET) {
                cft_set_error("instruction %lu closes a quiet region opened outside the loop "
                              "opened at instruction %lu", (unsigned long)pc,
                              (unsigned long)bpc[nb - 1]);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            nb--;
            qdepth--;
        } else if ((d.op == SEQ_ACTALL || d.op == SEQ_HALT) && depth > 0) {
            cft_set_error("instruction %lu is %s inside a loop, which would make the "
                          "all-lanes-done early exit observable", (unsigned long)pc,
                          ctrl_name(d.op));
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (d.op == SEQ_HALT && qdepth > 0) {
            cft_set_error("instruction %lu is HALT inside the quiet region opened at "
                          "instruction %lu", (unsigned long)pc, (unsigned long)bpc[nb - 1]);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (worst > kMaxIssue)
            return worst_refusal(worst, pc, false);
    }
    if (depth != 0) {
        cft_set_error("the program ends inside %d open loop%s", depth, depth == 1 ? "" : "s");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (qdepth != 0) {
        cft_set_error("the program ends inside %d open quiet region%s", qdepth,
                      qdepth == 1 ? "" : "s");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (worst > kMaxIssue)
        return worst_refusal(worst, p->n_insns - 1u, true);
    return CFT_OK;
}

// seq_check_against_device (program.c:959-1100) for this device's features:
// every one the reference defines but augmentedAddition.
cft_status check_against_device(const cft_program *p)
{
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const Dec d = decode(p->insns[pc]);
        if (d.ctrl && (d.op == SEQ_AUGADD || d.op == SEQ_AUGERR)) {
            cft_set_error("instruction %lu is %s (augmentedAddition) and this device does "
                          "not publish CFT_SEQ_FEAT_AUGADD (cft_caps.seq_features bit 15)",
                          (unsigned long)pc, ctrl_name(d.op));
            return CFT_ERR_UNSUPPORTED;
        }
    }
    return CFT_OK;
}

// seq_scratch_footprint (program.c:1103-1140)
void scratch_footprint(cft_program *p)
{
    uint32_t used = 0;
    bool indexed = false, touched = false;
    p->log2depth = 0;
    while ((1u << p->log2depth) < p->depth)
        p->log2depth++;
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const Dec d = decode(p->insns[pc]);
        if (!d.ctrl)
            continue;
        if (d.op == SEQ_STL || d.op == SEQ_LDL) {
            touched = true;
            used = std::max(used, (d.imm & kKxIndices) + 1u);
        } else if (d.op == SEQ_STX || d.op == SEQ_LDX) {
            touched = indexed = true;
        }
    }
    p->scratch_used = indexed ? p->depth : used;
    p->scratch_active = touched || (p->flags & CFT_PROG_FLAG_SCRATCH_IO);
}

// The slots a run clears and the pool holds (host/src/seq_exec.h:716-733).
uint32_t scratch_reach(const cft_program *p)
{
    if (!p->scratch_active)
        return 0;
    uint32_t r = std::max({p->scratch_used, p->n_in, p->n_out});
    return std::min(r, p->depth);
}

// The decoded form: registers renamed densely, loops matched.
void predecode(cft_program *p)
{
    bool used[32] = {false};
    std::vector<Dec> ds(p->n_insns);
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const Dec d = ds[pc] = decode(p->insns[pc]);
        if (!d.ctrl) {
            used[reg5(d.rd, d.hd)] = true;
            if (!d.ka) used[reg5(d.ra, d.ha)] = true;
            if (!d.kb) used[reg5(d.rb, d.hb)] = true;
            if (!d.kc) used[reg5(d.rc, d.hc)] = true;
            continue;
        }
        switch (d.op) {
        case SEQ_DEPOSIT: case SEQ_SETACT: case SEQ_STL: case SEQ_RAISE:
            used[reg5(d.ra, d.ha)] = true;
            break;
        case SEQ_LDL:
            used[reg5(d.rd, d.hd)] = true;
            break;
        case SEQ_STX:
            used[reg5(d.ra, d.ha)] = used[reg5(d.rb, d.hb)] = true;
            break;
        case SEQ_LDX:
            used[reg5(d.rd, d.hd)] = used[reg5(d.rb, d.hb)] = true;
            break;
        default:
            break;
        }
    }
    int phys[32];
    p->nregs = 0;
    for (int r = 0; r < 32; r++)
        phys[r] = used[r] ? (int)p->nregs++ : -1;
    for (int s = 0; s < 3; s++)
        p->phys_abc[s] = phys[s];
    if (p->nregs == 0)
        p->nregs = 1;

    p->code.assign(p->n_insns, SeqInsn{0, 0, 0, 0});
    std::vector<uint32_t> open;
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const Dec &d = ds[pc];
        SeqInsn &o = p->code[pc];
        auto ph = [&](int lo, int hi) { return (uint32_t)std::max(phys[reg5(lo, hi)], 0); };
        o.w = (uint32_t)d.op | ((uint32_t)d.ctrl << 8) | ((uint32_t)d.rnd << 9);
        o.imm = d.imm;
        if (!d.ctrl) {
            const uint32_t a = d.ka ? konst_index(d, 0) : ph(d.ra, d.ha);
            const uint32_t b = d.kb ? konst_index(d, 1) : ph(d.rb, d.hb);
            const uint32_t c = d.kc ? konst_index(d, 2) : ph(d.rc, d.hc);
            o.w |= ((uint32_t)d.ka << 12) | ((uint32_t)d.kb << 13) | ((uint32_t)d.kc << 14) |
                   (ph(d.rd, d.hd) << 16);
            o.src = a | (b << 9) | (c << 18);
            continue;
        }
        switch (d.op) {
        case SEQ_REPEAT:
            if (open.empty() && p->top_repeat == 0xFFFFFFFFu)
                p->top_repeat = pc;
            open.push_back(pc);
            break;
        case SEQ_ENDREP:
            o.aux = open.back() + 1;
            p->code[open.back()].aux = pc;
            open.pop_back();
            break;
        case SEQ_DEPOSIT: case SEQ_SETACT: case SEQ_STL: case SEQ_RAISE:
            o.src = ph(d.ra, d.ha);
            break;
        case SEQ_LDL:
            o.w |= ph(d.rd, d.hd) << 16;
            break;
        case SEQ_STX:
            o.src = ph(d.ra, d.ha) | (ph(d.rb, d.hb) << 9);
            break;
        case SEQ_LDX:
            o.w |= ph(d.rd, d.hd) << 16;
            o.src = ph(d.rb, d.hb) << 9;
            break;
        default:
            break;
        }
    }
}

void free_device(cft_program *p)
{
    if (p->d_code || p->d_consts || p->gexec || !p->wide.empty()) {
        if (hipSetDevice(p->dev->hip_index) == hipSuccess) {
            (void)hipStreamSynchronize(p->dev->stream);
            seq_drop_graph(p);
            if (p->d_code) (void)hipFree(p->d_code);
            if (p->d_consts) (void)hipFree(p->d_consts);
            for (auto &kv : p->wide)
                if (kv.second.d_recs) (void)hipFree(kv.second.d_recs);
        }
        (void)hipGetLastError();
    }
    p->d_code = nullptr;
    p->d_consts = nullptr;
    p->wide.clear();
}

cft_status upload(cft_program *p, const char *who)
{
    cft_status st = dev_enter(p->dev, who);
    if (st != CFT_OK)
        return st;
    const size_t cb = std::max<size_t>(1, p->code.size()) * sizeof(SeqInsn);
    hipError_t e = hipMalloc(&p->d_code, cb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p->d_code = nullptr;
        cft_set_error("%s: %zu bytes of device memory for the program could not be "
                      "allocated", who, cb);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    if (!p->code.empty())
        e = hipMemcpy(p->d_code, p->code.data(), p->code.size() * sizeof(SeqInsn),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && !(p->flags & CFT_PROG_FLAG_BANK_EXT) && p->n_consts) {
        const size_t kb = (size_t)p->n_consts * p->esz;
        e = hipMalloc(&p->d_consts, kb);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p->d_consts = nullptr;
            cft_set_error("%s: %zu bytes of device memory for the constants could not be "
                          "allocated", who, kb);
            return CFT_ERR_OUT_OF_MEMORY;
        }
        e = hipMemcpy(p->d_consts, p->image.data() + kHeader, kb, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
        return dev_fail(p->dev, e, who, "uploading the program");
    return CFT_OK;
}

} // namespace

extern "C" CFT_API void cft_program_free(cft_program *prog)
{
    if (!prog)
        return;
    free_device(prog);
    delete prog;
}

namespace {

// Everything cft_program_load checks, in the reference's order, against a
// scratch depth (the device's max_scratch).  Fills prog's header fields and
// its instruction words; allocates nothing on a device.
cft_status parse_image(const uint8_t *p, size_t bytes, uint32_t depth, cft_program *prog)
{
    if (bytes < kHeader) {
        cft_set_error("this image is %zu bytes, shorter than the %zu-byte header", bytes,
                      kHeader);
        return CFT_ERR_ARTIFACT;
    }
    const uint32_t magic = rd32(p), ver = rd32(p + 4), n_insns = rd32(p + 8),
                   n_consts = rd32(p + 12), maxdep = rd32(p + 16), prec = rd32(p + 20),
                   flags = rd32(p + 24), scr_io = rd32(p + 28);
    if (magic != kMagic) {
        cft_set_error("this image's magic is 0x%08lx, not \"CFTP\" (0x%08lx)",
                      (unsigned long)magic, (unsigned long)kMagic);
        return CFT_ERR_ARTIFACT;
    }
    if (ver != kVersion) {
        cft_set_error("this image is program-image version %lu; this library reads version "
                      "%lu", (unsigned long)ver, (unsigned long)kVersion);
        return CFT_ERR_ARTIFACT;
    }
    if (flags & ~kFlagsKnown) {
        cft_set_error("this image's header flags 0x%08lx carry unknown bits 0x%08lx",
                      (unsigned long)flags, (unsigned long)(flags & ~kFlagsKnown));
        return CFT_ERR_ARTIFACT;
    }
    if (!(flags & CFT_PROG_FLAG_SCRATCH_IO) && scr_io) {
        cft_set_error("this image's scratch_io word is 0x%08lx without SCRATCH_IO; it must "
                      "be zero", (unsigned long)scr_io);
        return CFT_ERR_ARTIFACT;
    }
    if (prec > 3) {
        cft_set_error("this image's precision code is %lu; the formats are 0 to 3",
                      (unsigned long)prec);
        return CFT_ERR_ARTIFACT;
    }
    const size_t esz = cft_format_size((cft_format)prec);
    const uint32_t n_stored = (flags & CFT_PROG_FLAG_BANK_EXT) ? 0u : n_consts;
    const uint64_t want = (uint64_t)kHeader + (uint64_t)n_stored * esz + (uint64_t)n_insns * 8;
    if ((uint64_t)bytes != want) {
        cft_set_error("this image is %llu bytes and its header describes %llu",
                      (unsigned long long)bytes, (unsigned long long)want);
        return CFT_ERR_ARTIFACT;
    }
    if (maxdep > cftl::kMaxDeposits)
        return cap_refusal("max_deposits", maxdep, cftl::kMaxDeposits, "deposit slots a lane",
                           "max_deposits");
    if (n_consts > cftl::kMaxConsts) {
        cft_set_error("this image declares %lu constants and an instruction addresses at "
                      "most %lu", (unsigned long)n_consts, (unsigned long)cftl::kMaxConsts);
        return CFT_ERR_UNSUPPORTED;
    }
    uint32_t n_in = 0, n_out = 0;
    if (flags & CFT_PROG_FLAG_SCRATCH_IO) {
        n_in = scr_io & 0xFFFFu;
        n_out = scr_io >> 16;
        if (n_in > depth)
            return cap_refusal("n_scratch_in", n_in, depth, "scratch slots a lane",
                               "max_scratch");
        if (n_out > depth)
            return cap_refusal("n_scratch_out", n_out, depth, "scratch slots a lane",
                               "max_scratch");
    }
    prog->fmt = (int)prec;
    prog->esz = esz;
    prog->max_deposits = maxdep;
    prog->n_insns = n_insns;
    prog->n_consts = n_consts;
    prog->flags = flags;
    prog->n_in = n_in;
    prog->n_out = n_out;
    prog->depth = depth;
    try {
        prog->image.assign(p, p + bytes);
        prog->insns.resize(n_insns);
    } catch (...) {
        return CFT_ERR_OUT_OF_MEMORY;
    }
    const uint8_t *ip = p + kHeader + (size_t)n_stored * esz;
    for (uint32_t i = 0; i < n_insns; i++)
        prog->insns[i] = (uint64_t)rd32(ip + 8 * (size_t)i) |
                         ((uint64_t)rd32(ip + 8 * (size_t)i + 4) << 32);

    cft_status st = validate(prog);
    if (st == CFT_OK && n_insns > cftl::kMaxInsns)
        st = cap_refusal("instruction count", n_insns, cftl::kMaxInsns,
                         "instructions it can hold", "max_insns");
    if (st == CFT_OK)
        st = check_against_device(prog);
    if (st != CFT_OK)
        return st;
    scratch_footprint(prog);
    predecode(prog);
    return CFT_OK;
}

} // namespace

extern "C" CFT_API cft_status cft_program_load(cft_device *dev, const void *image,
                                               size_t bytes, cft_program **out)
{
    static const char *who = "cft_program_load";
    cft_clear_error();
    if (!dev || !image || !out) {
        cft_set_error("cft_program_load was given a NULL %s",
                      !dev ? "device" : !image ? "image" : "out");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    cft_status st = host_in(dev, image, who);
    if (st != CFT_OK)
        return st;
    cft_program *prog = new (std::nothrow) cft_program();
    if (!prog)
        return CFT_ERR_OUT_OF_MEMORY;
    prog->dev = dev;
    st = parse_image((const uint8_t *)image, bytes, dev->scratch_depth, prog);
    if (st == CFT_OK)
        st = upload(prog, who);
    if (st != CFT_OK) {
        cft_program_free(prog);
        return st;
    }
    *out = prog;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_program_check(const void *image, size_t bytes,
                                                 uint32_t scratch_depth,
                                                 cft_program_info *info)
{
    cft_clear_error();
    if (!image) {
        cft_set_error("cftx_program_check was given a NULL image");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (scratch_depth == 0)
        scratch_depth = cftl::kDefaultDepth;
    if ((scratch_depth & (scratch_depth - 1)) || scratch_depth > cftl::kMaxDepth) {
        cft_set_error("cftx_program_check: scratch_depth %lu is not a power of two up to %lu",
                      (unsigned long)scratch_depth, (unsigned long)cftl::kMaxDepth);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_program prog;
    const cft_status st = parse_image((const uint8_t *)image, bytes, scratch_depth, &prog);
    if (st == CFT_OK && info && info->struct_size >= sizeof(size_t)) {
        cft_program_info v;
        memset(&v, 0, sizeof v);
        v.format = (cft_format)prog.fmt;
        v.max_deposits = prog.max_deposits;
        v.n_insns = prog.n_insns;
        v.n_consts = prog.n_consts;
        v.flags = prog.flags;
        v.n_scratch_in = prog.n_in;
        v.n_scratch_out = prog.n_out;
        v.scratch_used = prog.scratch_used;
        const size_t want = std::min(info->struct_size, sizeof v);
        v.struct_size = want;
        memcpy(info, &v, want);
    }
    return st;
}

extern "C" CFT_API cft_status cftx_program_wide(const void *image, size_t bytes,
                                                uint32_t scratch_depth, uint32_t width,
                                                uint32_t max_vregs, void *records, size_t cap,
                                                uint32_t info[6])
{
    cft_clear_error();
    if (!image || (!records && cap)) {
        cft_set_error("cftx_program_wide was given a NULL %s", !image ? "image" : "records");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (width == 0 || width > 64 || (width & (width - 1))) {
        cft_set_error("cftx_program_wide: width %lu is not a power of two up to 64",
                      (unsigned long)width);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (scratch_depth == 0)
        scratch_depth = cftl::kDefaultDepth;
    if ((scratch_depth & (scratch_depth - 1)) || scratch_depth > cftl::kMaxDepth) {
        cft_set_error("cftx_program_wide: scratch_depth %lu is not a power of two up to %lu",
                      (unsigned long)scratch_depth, (unsigned long)cftl::kMaxDepth);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_program prog;
    cft_status st = parse_image((const uint8_t *)image, bytes, scratch_depth, &prog);
    if (st != CFT_OK)
        return st;
    WideProgram wp;
    if (!seq_wide_eligible(&prog) ||
        !seq_wide_compile(&prog, (int)width, max_vregs ? std::min<uint32_t>(max_vregs, 65535)
                                                       : 65535, &wp)) {
        cft_set_error("cftx_program_wide: the program does not run wide (it deposits, sets "
                      "activity, indexes scratch or controls flags, or does not schedule)");
        return CFT_ERR_UNSUPPORTED;
    }
    const size_t want = (size_t)wp.nsteps * width;
    if (records)
        memcpy(records, wp.recs.data(), std::min(cap, want) * sizeof(WideRec));
    if (info) {
        info[0] = wp.nsteps;
        info[1] = wp.nvregs;
        info[2] = wp.nregs;
        info[3] = wp.top_step;
        info[4] = wp.body_steps;
        info[5] = wp.body_alu;
    }
    return CFT_OK;
}

extern "C" CFT_API cft_status cft_program_get_info(cft_program *prog, cft_program_info *out)
{
    cft_clear_error();
    if (!prog || !out)
        return CFT_ERR_INVALID_ARGUMENT;
    size_t want = out->struct_size;
    if (want < sizeof(size_t))
        return CFT_ERR_INVALID_ARGUMENT;
    cft_program_info info;
    memset(&info, 0, sizeof info);
    info.format = (cft_format)prog->fmt;
    info.max_deposits = prog->max_deposits;
    info.n_insns = prog->n_insns;
    info.n_consts = prog->n_consts;
    info.flags = prog->flags;
    info.n_scratch_in = prog->n_in;
    info.n_scratch_out = prog->n_out;
    info.scratch_used = prog->scratch_used;
    want = std::min(want, sizeof info);
    info.struct_size = want;
    memcpy(out, &info, want);
    return CFT_OK;
}

namespace {

// seq_check_bank (program.c:1640-1680)
cft_status check_bank(const cft_program *p, const void *bank, size_t bank_bytes,
                      const char *who)
{
    if (!(p->flags & CFT_PROG_FLAG_BANK_EXT)) {
        if (bank || bank_bytes) {
            cft_set_error("%s was given a %zu-byte constant bank and this program carries "
                          "its own %lu constants", who, bank_bytes,
                          (unsigned long)p->n_consts);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        return CFT_OK;
    }
    const size_t want = (size_t)p->n_consts * p->esz;
    if (bank_bytes != want) {
        cft_set_error("%s was given a %zu-byte constant bank and this program's is %zu bytes "
                      "(%lu %s constants)", who, bank_bytes, want, (unsigned long)p->n_consts,
                      cft_format_name((cft_format)p->fmt));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (want && !bank) {
        cft_set_error("%s: this program's constants arrive with the run and the bank "
                      "pointer is NULL", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

// seq_check_scratch (program.c:1695-1770)
cft_status check_scratch(const cft_program *p, const cft_run_args *A, const char *who)
{
    struct Side { const void *buf; size_t bytes; uint32_t slots; const char *name; };
    const Side side[2] = {{A->scratch_in, A->scratch_in_bytes, p->n_in, "in"},
                          {A->scratch_out, A->scratch_out_bytes, p->n_out, "out"}};
    if (!(p->flags & CFT_PROG_FLAG_SCRATCH_IO)) {
        for (int w = 0; w < 2; w++)
            if (side[w].buf || side[w].bytes) {
                cft_set_error("%s was given a %zu-byte scratch-%s block and this program "
                              "declares no scratch I/O", who, side[w].bytes, side[w].name);
                return CFT_ERR_INVALID_ARGUMENT;
            }
        return CFT_OK;
    }
    for (int w = 0; w < 2; w++) {
        size_t want;
        const size_t slots = side[w].slots;
        if (!slots || !A->n)
            want = 0;
        else if (A->n > SIZE_MAX / slots / p->esz)
            want = SIZE_MAX;
        else
            want = A->n * slots * p->esz;
        if (w == 0 && A->idx_scratch_in && want != SIZE_MAX) {
            want = A->idx_scratch_src > SIZE_MAX / p->esz ? SIZE_MAX
                                                          : A->idx_scratch_src * p->esz;
            if (want != SIZE_MAX && side[0].bytes != want) {
                cft_set_error("%s was given a %zu-byte scratch-in pool and idx_scratch_src "
                              "names %zu elements (%zu bytes)", who, side[0].bytes,
                              A->idx_scratch_src, want);
                return CFT_ERR_INVALID_ARGUMENT;
            }
        }
        if (want == SIZE_MAX) {
            cft_set_error("%s: the scratch-%s block of %lu slots over %zu lanes is more than "
                          "this process can address", who, side[w].name,
                          (unsigned long)side[w].slots, A->n);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (side[w].bytes != want) {
            cft_set_error("%s was given a %zu-byte scratch-%s block and this program's is "
                          "%zu bytes (%zu lanes x %lu slots, lane-major)", who, side[w].bytes,
                          side[w].name, want, A->n, (unsigned long)side[w].slots);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (want && !side[w].buf) {
            cft_set_error("%s: this program declares %lu scratch-%s slots a lane and the "
                          "pointer is NULL", who, (unsigned long)side[w].slots, side[w].name);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    return CFT_OK;
}

// seq_check_round2 (program.c:2050-2250): index tables, lane mask, lane flags.
cft_status check_indexed_args(const cft_program *p, const cft_run_args *A, const char *who)
{
    const uint32_t *idx[3] = {A->idx_a, A->idx_b, A->idx_c};
    const size_t src[3] = {A->idx_a_src, A->idx_b_src, A->idx_c_src};
    const void *strm[3] = {A->a, A->b, A->c};
    for (int r = 0; r < 3; r++) {
        if (!idx[r]) {
            if (src[r]) {
                cft_set_error("%s: idx_%c_src = %zu with no index table", who, 'a' + r,
                              src[r]);
                return CFT_ERR_INVALID_ARGUMENT;
            }
            continue;
        }
        if (!strm[r]) {
            cft_set_error("%s: idx_%c indexes stream %c, which is NULL", who, 'a' + r,
                          'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (src[r] == 0) {
            cft_set_error("%s: idx_%c is set and idx_%c_src is zero", who, 'a' + r, 'a' + r);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    if (!A->idx_scratch_in && A->idx_scratch_src) {
        cft_set_error("%s: idx_scratch_src = %zu with no idx_scratch_in table", who,
                      A->idx_scratch_src);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (A->idx_scratch_in) {
        if (p->n_in == 0) {
            cft_set_error("%s: idx_scratch_in is set and this program declares no scratch "
                          "input", who);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (!A->scratch_in) {
            cft_set_error("%s: idx_scratch_in indexes scratch_in, which is NULL", who);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (A->idx_scratch_src == 0) {
            cft_set_error("%s: idx_scratch_in is set and idx_scratch_src is zero", who);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    if (!A->lane_mask && A->lane_mask_bytes) {
        cft_set_error("%s: lane_mask_bytes = %zu with no lane_mask", who, A->lane_mask_bytes);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (A->lane_mask && A->lane_mask_bytes != (A->n + 7u) / 8u) {
        cft_set_error("%s: lane_mask_bytes is %zu and a mask over %zu lanes is %zu bytes",
                      who, A->lane_mask_bytes, A->n, (A->n + 7u) / 8u);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!A->lane_flags && A->lane_flags_bytes) {
        cft_set_error("%s: lane_flags_bytes = %zu with no lane_flags", who,
                      A->lane_flags_bytes);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (A->lane_flags && A->lane_flags_bytes != A->n) {
        cft_set_error("%s: lane_flags_bytes is %zu and the lane flags of %zu lanes are %zu "
                      "bytes", who, A->lane_flags_bytes, A->n, A->n);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (A->n) {
        for (int r = 0; r < 3; r++) {
            cft_status hs = host_in(p->dev, idx[r], who);
            if (hs != CFT_OK)
                return hs;
        }
        cft_status hs = host_in(p->dev, A->idx_scratch_in, who);
        if (hs != CFT_OK)
            return hs;
    }
    for (int r = 0; r < 3; r++) {
        if (!idx[r])
            continue;
        for (size_t i = 0; i < A->n; i++)
            if (idx[r][i] != CFT_IDX_NONE && (size_t)idx[r][i] >= src[r]) {
                cft_set_error("%s: idx_%c[%zu] = %lu is at or past the %zu elements of "
                              "stream %c", who, 'a' + r, i, (unsigned long)idx[r][i], src[r],
                              'a' + r);
                return CFT_ERR_INVALID_ARGUMENT;
            }
    }
    if (A->idx_scratch_in) {
        const size_t entries = A->n * (size_t)p->n_in;
        for (size_t e = 0; e < entries; e++)
            if (A->idx_scratch_in[e] != CFT_IDX_NONE &&
                (size_t)A->idx_scratch_in[e] >= A->idx_scratch_src) {
                cft_set_error("%s: idx_scratch_in[%zu] = %lu (lane %zu slot %zu) is at or "
                              "past the %zu elements of the pool", who, e,
                              (unsigned long)A->idx_scratch_in[e], e / p->n_in, e % p->n_in,
                              A->idx_scratch_src);
                return CFT_ERR_INVALID_ARGUMENT;
            }
    }
    return CFT_OK;
}

} // namespace

Shape shape_for(const cft_program *p, int waves_pref)
{
    const size_t cap = p->dev->lds_max;
    const bool code_fits = (size_t)p->n_insns * 16 <= 32768;
    for (int pass = 0; pass < 2; pass++) {
        const bool cil = pass == 0 && code_fits;
        if (pass == 0 && !code_fits)
            continue;
        for (int w = waves_pref; w >= 1; w /= 2)
            if (cftk_seq_lds(p->fmt, cil ? p->n_insns : 0, p->n_consts, p->nregs, w) <= cap)
                return {w, cil};
    }
    return {0, false};
}


// The run itself: everything after the argument rules.  trips: per-lane trip
// counts of the program's first top-level REPEAT (device pointer), or NULL.
cft_status seq_execute(cft_program *prog, const cft_run_args *A, const uint32_t *d_trips,
                       int waves_pref, const char *who)
{
    cft_device *dev = prog->dev;
    const size_t n = A->n;
    if (n == 0) {
        cft_emit_flags(dev, 0, A->flags_out);
        return CFT_OK;
    }
    cft_status st = host_in(dev, A->bank, who);
    if (st != CFT_OK)
        return st;
    st = dev_enter(dev, who);
    if (st != CFT_OK)
        return st;
    const Shape sh = shape_for(prog, waves_pref);
    if (sh.waves == 0) {
        cft_set_error("%s: this program's %lu registers and %lu constants do not fit one "
                      "workgroup's %zu bytes of LDS", who, (unsigned long)prog->nregs,
                      (unsigned long)prog->n_consts, dev->lds_max);
        return CFT_ERR_UNSUPPORTED;
    }
    const size_t esz = prog->esz;
    const bool masked = A->lane_mask != nullptr;
    SeqArgs S;
    memset(&S, 0, sizeof S);

    // inputs
    const void *da = nullptr, *db = nullptr, *dc = nullptr, *dia = nullptr, *dib = nullptr,
               *dic = nullptr, *disi = nullptr, *dsi = nullptr, *dbank = nullptr,
               *dmask = nullptr;
    const size_t sa = (A->idx_a ? A->idx_a_src : n) * esz;
    const size_t sb = (A->idx_b ? A->idx_b_src : n) * esz;
    const size_t sc = (A->idx_c ? A->idx_c_src : n) * esz;
    if ((st = bind_in(dev, ROLE_A, A->a, sa, &da, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_B, A->b, sb, &db, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_C, A->c, sc, &dc, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_IA, A->idx_a, n * 4, &dia, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_IB, A->idx_b, n * 4, &dib, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_IC, A->idx_c, n * 4, &dic, who)) != CFT_OK ||
        (st = bind_in(dev, ROLE_MASK, A->lane_mask, A->lane_mask_bytes, &dmask, who)) !=
            CFT_OK)
        return st;
    if (prog->n_in && A->scratch_in) {
        const size_t sib = A->scratch_in_bytes;
        if ((st = bind_in(dev, ROLE_SIN, A->scratch_in, sib, &dsi, who)) != CFT_OK ||
            (st = bind_in(dev, ROLE_ISI, A->idx_scratch_in, n * prog->n_in * 4, &disi, who)) !=
                CFT_OK)
            return st;
    }
    if (prog->flags & CFT_PROG_FLAG_BANK_EXT) {
        if (A->bank_bytes &&
            (st = bind_in(dev, ROLE_BANK, A->bank, A->bank_bytes, &dbank, who)) != CFT_OK)
            return st;
    } else {
        dbank = prog->d_consts;
    }

    // outputs: a staged window is pre-filled from the caller's array where a
    // masked lane leaves it untouched, so the copy back returns it unchanged
    struct Out { int role; void *host; size_t bytes; void *dptr; cft_buffer *buf; };
    Out outs[4] = {
        {ROLE_D, A->deposits, prog->max_deposits ? n * prog->max_deposits * esz : 0, nullptr,
         nullptr},
        {ROLE_COUNTS, A->counts, A->counts ? n * 4 : 0, nullptr, nullptr},
        {ROLE_LFLAGS, A->lane_flags, A->lane_flags ? n : 0, nullptr, nullptr},
        {ROLE_SOUT, A->scratch_out, prog->n_out ? n * prog->n_out * esz : 0, nullptr, nullptr},
    };
    for (Out &o : outs) {
        if (!o.host || !o.bytes)
            continue;
        if ((st = bind_out(dev, o.role, o.host, o.bytes, &o.dptr, &o.buf, who)) != CFT_OK)
            return st;
        if (!o.buf && masked) {
            const hipError_t e =
                hipMemcpyAsync(o.dptr, o.host, o.bytes, hipMemcpyHostToDevice, dev->stream);
            if (e != hipSuccess)
                return dev_fail(dev, e, who, "staging an output window");
        }
    }

    // the scratch pool, in chunks of lanes when it would not fit at once
    const uint32_t reach = scratch_reach(prog);
    const size_t per_lane = (size_t)reach * esz;
    const size_t lanes_per_block = (size_t)sh.waves * 64;
    size_t chunk = n;
    if (per_lane) {
        const size_t budget = (size_t)8 << 30;
        chunk = std::min(n, std::max<size_t>(lanes_per_block, budget / per_lane));
        chunk = (chunk + lanes_per_block - 1) / lanes_per_block * lanes_per_block;
        if ((st = arena_reserve(dev, dev->stage[ROLE_POOL], chunk * per_lane, who)) != CFT_OK)
            return st;
    }

    S.code = prog->d_code;
    S.n_insns = prog->n_insns;
    S.n_consts = prog->n_consts;
    S.bank = (const uint32_t *)dbank;
    S.a = (const uint32_t *)da;
    S.b = (const uint32_t *)db;
    S.c = (const uint32_t *)dc;
    S.idx_a = (const uint32_t *)dia;
    S.idx_b = (const uint32_t *)dib;
    S.idx_c = (const uint32_t *)dic;
    S.sin = (const uint32_t *)dsi;
    S.idx_si = (const uint32_t *)disi;
    S.n_in = prog->n_in;
    S.n_out = prog->n_out;
    S.sout = (uint32_t *)outs[3].dptr;
    S.pool = per_lane ? (uint32_t *)dev->stage[ROLE_POOL].p : nullptr;
    S.npool = chunk;
    S.reach = reach;
    S.depth = prog->depth;
    S.log2depth = prog->log2depth;
    S.strict = (prog->flags & CFT_PROG_FLAG_SCRATCH_STRICT) ? 1 : 0;
    S.dep = (uint32_t *)outs[0].dptr;
    S.max_dep = S.dep ? prog->max_deposits : 0;
    S.counts = (uint32_t *)outs[1].dptr;
    S.mask = (const uint8_t *)dmask;
    S.lflags = (uint8_t *)outs[2].dptr;
    S.words = dev->dwords;
    S.code_in_lds = sh.code_in_lds ? 1 : 0;
    S.trips = d_trips;

    hipError_t e = cftk_fill32(dev->dwords, 0u, 2, dev->stream);
    if (e == hipSuccess)
        e = timer_begin(dev);
    for (size_t lo = 0; e == hipSuccess && lo < n; lo += chunk) {
        S.lane0 = lo;
        S.n = std::min(n, lo + chunk);
        e = cftk_seq(prog->fmt, &S, prog->nregs, prog->phys_abc[0], prog->phys_abc[1],
                     prog->phys_abc[2], d_trips ? prog->top_repeat : 0xFFFFFFFFu, sh.waves,
                     dev->stream);
    }
    if (e == hipSuccess)
        e = timer_end(dev);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "launching the interpreter");
    for (Out &o : outs) {
        if (!o.host || !o.bytes || o.buf)
            continue;
        e = hipMemcpyAsync(o.host, o.dptr, o.bytes, hipMemcpyDeviceToHost, dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "copying an output back");
    }
    e = hipMemcpyAsync(dev->hwords, dev->dwords, 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "running the program");
    timer_collect(dev);
    for (Out &o : outs)
        if (o.buf)
            o.buf->host_valid = false;
    cft_emit_flags(dev, dev->hwords[0], A->flags_out);
    if (A->bus_out)
        *A->bus_out = dev->hwords[1];
    return CFT_OK;
}

// cft_program_run_ex's checks (program.c:2258-2330); also the batched
// runner's, which passes its own `who`.
cft_status seq_check_run(cft_program *prog, const cft_run_args *args, const char *who)
{
    if (args->n != 0) {
        const size_t esz = prog->esz;
        if (!args->a) {
            cft_set_error("%s: stream a is NULL; a run needs it (it initialises r0)", who);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if (!args->deposits && prog->max_deposits) {
            cft_set_error("%s: deposits is NULL for a program with max_deposits = %u", who,
                          (unsigned)prog->max_deposits);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        const size_t per = esz * (prog->max_deposits ? prog->max_deposits : 1u);
        if (args->n > SIZE_MAX / per) {
            cft_set_error("%s: n = %zu lanes is more than this library can size", who,
                          args->n);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
    cft_status st = check_bank(prog, args->bank, args->bank_bytes, who);
    if (st == CFT_OK)
        st = check_scratch(prog, args, who);
    if (st == CFT_OK)
        st = check_indexed_args(prog, args, who);
    return st;
}

extern "C" CFT_API cft_status cft_program_run_ex(cft_program *prog, const cft_run_args *args)
{
    static const char *who = "cft_program_run_ex";
    cft_clear_error();
    if (!prog || !args)
        return CFT_ERR_INVALID_ARGUMENT;
    if (args->struct_size != sizeof(cft_run_args)) {
        cft_set_error("cft_program_run_ex was given a %zu-byte cft_run_args and this "
                      "library's is %zu bytes", args->struct_size, sizeof(cft_run_args));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_run_args A = *args;
    if (A.bus_out)
        *A.bus_out = 0;
    cft_status st = seq_check_run(prog, &A, who);
    if (st != CFT_OK)
        return st;
    if (!A.bank_bytes)
        A.bank = nullptr;
    return seq_execute(prog, &A, nullptr, 4, who);
}

namespace {

void args_init(cft_run_args *A, const void *a, const void *b, const void *c, void *deposits,
               uint32_t *counts, size_t n, uint32_t *flags, uint32_t *bus)
{
    memset(A, 0, sizeof *A);
    A->struct_size = sizeof *A;
    A->a = a;
    A->b = b;
    A->c = c;
    A->n = n;
    A->deposits = deposits;
    A->counts = counts;
    A->flags_out = flags;
    A->bus_out = bus;
}

cft_status refuse_scratch_io(const cft_program *p, const char *who)
{
    if (!(p->flags & CFT_PROG_FLAG_SCRATCH_IO))
        return CFT_OK;
    cft_set_error("this program declares a scratch block (%lu in, %lu out) and %s has "
                  "nowhere to put it; call cft_program_run_ex", (unsigned long)p->n_in,
                  (unsigned long)p->n_out, who);
    return CFT_ERR_INVALID_ARGUMENT;
}

} // namespace

extern "C" CFT_API cft_status cft_program_run(cft_program *prog, const void *a, const void *b,
                                              const void *c, void *deposits, uint32_t *counts,
                                              size_t n, uint32_t *flags, uint32_t *bus)
{
    cft_clear_error();
    if (bus)
        *bus = 0;
    if (!prog)
        return CFT_ERR_INVALID_ARGUMENT;
    if (prog->flags & CFT_PROG_FLAG_BANK_EXT) {
        cft_set_error("this program's %lu constants arrive with each run (BANK_EXT); call "
                      "cft_program_run_bank", (unsigned long)prog->n_consts);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_status st = refuse_scratch_io(prog, "cft_program_run");
    if (st != CFT_OK)
        return st;
    cft_run_args A;
    args_init(&A, a, b, c, deposits, counts, n, flags, bus);
    return cft_program_run_ex(prog, &A);
}

extern "C" CFT_API cft_status cft_program_run_bank(cft_program *prog, const void *bank,
                                                   size_t bank_bytes, const void *a,
                                                   const void *b, const void *c,
                                                   void *deposits, uint32_t *counts, size_t n,
                                                   uint32_t *flags_out, uint32_t *bus_out)
{
    cft_clear_error();
    if (bus_out)
        *bus_out = 0;
    if (!prog)
        return CFT_ERR_INVALID_ARGUMENT;
    cft_status st = check_bank(prog, bank, bank_bytes, "cft_program_run_bank");
    if (st == CFT_OK)
        st = refuse_scratch_io(prog, "cft_program_run_bank");
    if (st != CFT_OK)
        return st;
    cft_run_args A;
    args_init(&A, a, b, c, deposits, counts, n, flags_out, bus_out);
    A.bank = bank;
    A.bank_bytes = bank_bytes;
    return cft_program_run_ex(prog, &A);
}

// program.c:2425-2454: SHA-256 over the image bytes, then the bank's.
extern "C" CFT_API cft_status cft_program_digest(cft_program *prog, const void *bank,
                                                 size_t bank_bytes, uint8_t out[32])
{
    cft_clear_error();
    if (!prog || !out)
        return CFT_ERR_INVALID_ARGUMENT;
    cft_status st = check_bank(prog, bank, bank_bytes, "cft_program_digest");
    if (st == CFT_OK)
        st = host_in(prog->dev, bank, "cft_program_digest");
    if (st != CFT_OK)
        return st;
    Sha256 s;
    s.update(prog->image.data(), prog->image.size());
    if (bank_bytes)
        s.update(bank, bank_bytes);
    s.final(out);
    return CFT_OK;
}

cft_status seq_check_bank(const cft_program *p, const void *bank, size_t bank_bytes,
                          const char *who)
{
    return check_bank(p, bank, bank_bytes, who);
}

uint32_t seq_scratch_reach(const cft_program *p) { return scratch_reach(p); }
