/* The oracle's sequencer executor: a CPU restatement of the reference's
 * program interpreter, lane by lane.
 *
 * TEST INFRASTRUCTURE ONLY - see sf_oracle.h.  It restates
 *   decode and operand sources   host/src/seq_exec.h:175-248
 *   the interpreter loop         host/src/seq_exec.h:412-685
 *   lanes in / out, deposits     host/src/seq_exec.h:699-923
 *   the software run's driver    host/src/program.c:1823-2028
 * for an image the loader has already accepted (the product's loader is
 * checked against the reference loader separately).
 *
 * One lane at a time.  The reference runs blocks of lanes SIMT, and its
 * contract (docs/SEQUENCER.md P2/P3) is that the block is invisible: a
 * lane's deposits are addressed by its own index, inactive lanes write and
 * raise nothing, and the only cross-lane effect - skipping a loop body no
 * lane is active in - cannot change a result.  Running a block of one lane
 * is therefore a restatement, not an approximation.
 *
 * Not restated: AUGADD/AUGERR (the MI355X path refuses them at load) and
 * the index tables of cft_run_args.
 */
#include "sf_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

enum { C_HALT, C_REPEAT, C_ENDREP, C_DEPOSIT, C_SETACT, C_ACTALL, C_STL,
       C_LDL, C_STX, C_LDX, C_AUGADD, C_AUGERR, C_QUIET, C_ENDQUIET,
       C_RAISE };

#define ST_DEPOSIT_OVERFLOW 0x10u
#define ST_SCRATCH_RANGE    0x20u
#define ST_MARKED           0x40u
#define PF_BANK_EXT         1u
#define PF_SCRATCH_IO       2u
#define PF_STRICT           4u

typedef struct {
    const oc_fmt *f;
    int fmt, esz;
    uint32_t n_insns, n_consts, max_dep, flags, n_in, n_out;
    const uint8_t *insn_bytes;
    oc_big *konst;
    uint32_t depth;
    int log2depth;
    /* run */
    const uint8_t *a, *b, *c, *scratch_in, *lane_mask;
    uint8_t *scratch_out, *deposits, *lane_flags;
    uint32_t *counts;
} prog_t;

static uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

static uint64_t insn_at(const prog_t *P, uint32_t pc)
{
    const uint8_t *q = P->insn_bytes + (size_t)pc * 8;
    return (uint64_t)rd32(q) | ((uint64_t)rd32(q + 4) << 32);
}

/* host/src/seq_exec.h:175-196 and 222-237: a field's register (five bits,
 * the fifth in imm[27:24]) or constant index (four bits, or under kx a byte
 * of imm plus its ninth bit in imm[30:28]). */
typedef struct {
    int op, rd, rnd, ctrl;
    int src[3], isk[3];
    uint32_t imm;
} dec_t;

static void decode(uint64_t w, dec_t *d)
{
    int k;
    int kx = (int)((w >> 30) & 1);
    d->op = (int)(w & 0xFF);
    d->rnd = (int)((w >> 24) & 7);
    d->ctrl = (int)((w >> 31) & 1);
    d->imm = (uint32_t)(w >> 32);
    d->rd = (int)((w >> 8) & 0xF) | (int)(((d->imm >> 24) & 1) << 4);
    for (k = 0; k < 3; k++) {
        int field = (int)((w >> (12 + 4 * k)) & 0xF);
        int hi = (int)((d->imm >> (25 + k)) & 1);
        d->isk[k] = (int)((w >> (27 + k)) & 1);
        if (!d->isk[k])
            d->src[k] = field | (hi << 4);
        else if (kx)
            d->src[k] = (int)((d->imm >> (8 * k)) & 0xFF) |
                        (int)(((d->imm >> (28 + k)) & 1) << 8);
        else
            d->src[k] = field;
    }
}

static int bitlen_big(const oc_big *a)
{
    uint32_t t;
    int b = 0;
    if (a->n == 0) return 0;
    t = a->w[a->n - 1];
    while (t) { b++; t >>= 1; }
    return (a->n - 1) * 32 + b;
}

/* the matching ENDREP of the REPEAT at pc (the loader proved nesting) */
static uint32_t matching_endrep(const prog_t *P, uint32_t pc)
{
    int depth = 0;
    uint32_t j;
    for (j = pc; j < P->n_insns; j++) {
        dec_t d;
        decode(insn_at(P, j), &d);
        if (!d.ctrl) continue;
        if (d.op == C_REPEAT) depth++;
        else if (d.op == C_ENDREP && --depth == 0) return j;
    }
    return P->n_insns;
}

typedef struct {
    oc_big regs[32];
    oc_big *scratch;           /* depth slots */
    int active, keep;
    uint32_t count;
    uint8_t lflags;
} lane_t;

/* One lane through the program: host/src/seq_exec.h:412-685 for a block of
 * one.  Returns 0, or 1 on an internal failure. */
static int run_lane(const prog_t *P, lane_t *L, size_t lane, uint32_t *flags,
                    uint32_t *status)
{
    struct { uint32_t body, left; } stk[4];
    int sp = 0, qdepth = 0;
    uint32_t pc = 0;
    const int wbits = P->f->width;

    while (pc < P->n_insns) {
        dec_t d;
        decode(insn_at(P, pc), &d);
        if (!d.ctrl) {
            const oc_big *s[3];
            oc_big out;
            uint32_t fl = 0;
            int k;
            if (L->active) {
                for (k = 0; k < 3; k++)
                    s[k] = d.isk[k] ? &P->konst[d.src[k]] : &L->regs[d.src[k]];
                oc_compute_big(P->f, d.op, d.rnd, s[0], s[1], s[2], &out, &fl);
                L->regs[d.rd] = out;
                if (!qdepth) {
                    *flags |= fl;
                    L->lflags |= (uint8_t)fl;
                }
            }
            pc++;
            continue;
        }
        switch (d.op) {
        case C_HALT:
            return 0;
        case C_REPEAT:
            if (d.imm == 0 || !L->active) {
                pc = matching_endrep(P, pc) + 1;
                break;
            }
            stk[sp].body = pc + 1;
            stk[sp].left = d.imm;
            sp++;
            pc++;
            break;
        case C_ENDREP:
            stk[sp - 1].left--;
            if (stk[sp - 1].left > 0 && L->active) {
                pc = stk[sp - 1].body;
            } else {
                sp--;
                pc++;
            }
            break;
        case C_DEPOSIT:
            if (L->active) {
                if (L->count >= P->max_dep) {
                    *status |= ST_DEPOSIT_OVERFLOW;
                    L->lflags |= 0x20;
                } else {
                    size_t slot = lane * P->max_dep + L->count;
                    oc_store(&L->regs[d.src[0]],
                             P->deposits + slot * (size_t)P->esz, P->esz);
                    L->count++;
                }
            }
            pc++;
            break;
        case C_SETACT:
            if (L->active) {
                oc_big m = L->regs[d.src[0]];
                int top = wbits - 1;
                if ((top >> 5) < m.n) {
                    m.w[top >> 5] &= ~(1u << (top & 31));
                    while (m.n > 0 && m.w[m.n - 1] == 0) m.n--;
                }
                L->active = m.n != 0;
            }
            pc++;
            break;
        case C_ACTALL:
            L->active = L->keep;
            pc++;
            break;
        case C_STL:
        case C_LDL: {
            uint32_t slot = d.imm & 0x00FFFFFFu;
            if (L->active) {
                if (slot >= P->depth)
                    return 1;
                if (d.op == C_STL)
                    L->scratch[slot] = L->regs[d.src[0]];
                else
                    L->regs[d.rd] = L->scratch[slot];
            }
            pc++;
            break;
        }
        case C_STX:
        case C_LDX: {
            int rb = d.src[1];
            int step = (int)(d.imm & 0xFFFu);
            if (step & 0x800) step -= 0x1000;
            if (L->active) {
                const oc_big *ix = &L->regs[rb];
                if ((P->flags & PF_STRICT) && bitlen_big(ix) > P->log2depth) {
                    *status |= ST_SCRATCH_RANGE;
                    L->lflags |= 0x40;
                    if (d.op == C_LDX)
                        L->regs[d.rd].n = 0;
                } else {
                    uint32_t slot = (ix->n ? ix->w[0] : 0) & (P->depth - 1u);
                    if (d.op == C_STX)
                        L->scratch[slot] = L->regs[d.src[0]];
                    else
                        L->regs[d.rd] = L->scratch[slot];
                }
                /* revision 8's post-step: rb += step modulo 2^width, except
                 * an LDX into its own index register, which keeps the load */
                if (step && !(d.op == C_LDX && d.rd == rb)) {
                    oc_big mag, zero, out;
                    uint32_t fl = 0;
                    memset(&mag, 0, sizeof mag);
                    mag.w[0] = (uint32_t)(step < 0 ? -step : step);
                    mag.n = 1;
                    zero.n = 0;
                    oc_compute_big(P->f, step > 0 ? 19 : 20, 0, &L->regs[rb],
                                   &mag, &zero, &out, &fl);
                    L->regs[rb] = out;
                }
            }
            pc++;
            break;
        }
        case C_QUIET:
            qdepth++;
            pc++;
            break;
        case C_ENDQUIET:
            qdepth--;
            pc++;
            break;
        case C_RAISE:
            if (L->active) {
                const oc_big *r = &L->regs[d.src[0]];
                uint32_t low = (r->n ? r->w[0] : 0) & 0xFFu;
                if (!qdepth) {
                    *flags |= low & 0x1Fu;
                    L->lflags |= (uint8_t)(low & 0x1Fu);
                }
                if (low & 0x80u) {
                    *status |= ST_MARKED;
                    L->lflags |= 0x80;
                }
            }
            pc++;
            break;
        default:
            return 1;                       /* AUGADD/AUGERR: not restated */
        }
    }
    return 0;
}

typedef struct {
    const prog_t *P;
    size_t lo, hi;
    uint32_t flags, status;
    int err;
} seq_job;

static void *seq_slice(void *arg)
{
    seq_job *J = (seq_job *)arg;
    const prog_t *P = J->P;
    lane_t *L = (lane_t *)calloc(1, sizeof *L);
    size_t i;
    const size_t esz = (size_t)P->esz;
    J->flags = J->status = 0;
    J->err = 0;
    if (!L) { J->err = 1; return NULL; }
    L->scratch = (oc_big *)calloc(P->depth, sizeof(oc_big));
    if (!L->scratch) { free(L); J->err = 1; return NULL; }
    for (i = J->lo; i < J->hi; i++) {
        uint32_t s;
        int r;
        L->keep = !P->lane_mask || ((P->lane_mask[i >> 3] >> (i & 7)) & 1u);
        if (!L->keep)
            continue;                   /* a masked lane writes nothing */
        for (r = 0; r < 32; r++)
            L->regs[r].n = 0;
        oc_load(&L->regs[0], P->a + i * esz, P->esz);
        if (P->b) oc_load(&L->regs[1], P->b + i * esz, P->esz);
        if (P->c) oc_load(&L->regs[2], P->c + i * esz, P->esz);
        for (s = 0; s < P->depth; s++)
            L->scratch[s].n = 0;
        for (s = 0; P->scratch_in && s < P->n_in; s++)
            oc_load(&L->scratch[s],
                    P->scratch_in + (i * P->n_in + s) * esz, P->esz);
        if (P->deposits && P->max_dep)
            memset(P->deposits + i * P->max_dep * esz, 0, P->max_dep * esz);
        L->active = 1;
        L->count = 0;
        L->lflags = 0;
        if (run_lane(P, L, i, &J->flags, &J->status)) { J->err = 1; break; }
        if (P->counts) P->counts[i] = L->count;
        if (P->lane_flags) P->lane_flags[i] = L->lflags;
        for (s = 0; P->scratch_out && s < P->n_out; s++)
            oc_store(&L->scratch[s],
                     P->scratch_out + (i * P->n_out + s) * esz, P->esz);
    }
    free(L->scratch);
    free(L);
    return NULL;
}

/* Run a loaded image over n lanes.  bank: the BANK_EXT constants (n_consts
 * format-width values) or NULL for an image that carries its own.  depth: the
 * scratch slots a lane (a power of two).  Buffers as cft_run_args has them.
 * Returns 0, or 1 on a malformed call or an internal failure. */
int oc_seq_run(const uint8_t *image, size_t image_bytes, const uint8_t *bank,
               const uint8_t *a, const uint8_t *b, const uint8_t *c, size_t n,
               const uint8_t *scratch_in, uint8_t *scratch_out,
               uint8_t *deposits, uint32_t *counts, const uint8_t *lane_mask,
               uint8_t *lane_flags, uint32_t depth, int threads,
               uint32_t *flags_out, uint32_t *status_out)
{
    prog_t P;
    seq_job jobs[256];
    pthread_t tid[256];
    uint32_t i, prec, scr_io, fl = 0, st = 0;
    size_t kbytes;
    int t, err = 0;

    if (image_bytes < 32 || rd32(image) != 0x50544643u || !a)
        return 1;
    memset(&P, 0, sizeof P);
    P.n_insns = rd32(image + 8);
    P.n_consts = rd32(image + 12);
    P.max_dep = rd32(image + 16);
    prec = rd32(image + 20);
    P.flags = rd32(image + 24);
    scr_io = rd32(image + 28);
    if (prec > 3 || !depth || (depth & (depth - 1)))
        return 1;
    P.fmt = (int)prec;
    P.f = &oc_formats[prec];
    P.esz = P.f->width / 8;
    if (P.flags & PF_SCRATCH_IO) {
        P.n_in = scr_io & 0xFFFFu;
        P.n_out = scr_io >> 16;
    }
    kbytes = (P.flags & PF_BANK_EXT) ? 0 : (size_t)P.n_consts * P.esz;
    if (image_bytes != 32 + kbytes + (size_t)P.n_insns * 8)
        return 1;
    P.insn_bytes = image + 32 + kbytes;
    P.konst = (oc_big *)calloc(P.n_consts ? P.n_consts : 1, sizeof(oc_big));
    if (!P.konst)
        return 1;
    for (i = 0; i < P.n_consts; i++) {
        const uint8_t *src = (P.flags & PF_BANK_EXT) ? bank : image + 32;
        if (!src) { free(P.konst); return 1; }
        oc_load(&P.konst[i], src + (size_t)i * P.esz, P.esz);
    }
    P.depth = depth;
    while ((1u << P.log2depth) < depth) P.log2depth++;
    P.a = a; P.b = b; P.c = c;
    P.scratch_in = scratch_in; P.scratch_out = scratch_out;
    P.deposits = deposits; P.counts = counts;
    P.lane_mask = lane_mask; P.lane_flags = lane_flags;

    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    for (t = 0; t < threads; t++) {
        jobs[t].P = &P;
        jobs[t].lo = n * (size_t)t / (size_t)threads;
        jobs[t].hi = n * (size_t)(t + 1) / (size_t)threads;
    }
    if (threads == 1) {
        seq_slice(&jobs[0]);
    } else {
        for (t = 0; t < threads; t++)
            pthread_create(&tid[t], NULL, seq_slice, &jobs[t]);
        for (t = 0; t < threads; t++)
            pthread_join(tid[t], NULL);
    }
    for (t = 0; t < threads; t++) {
        fl |= jobs[t].flags;
        st |= jobs[t].status;
        err |= jobs[t].err;
    }
    free(P.konst);
    if (flags_out) *flags_out = fl;
    if (status_out) *status_out = st;
    return err;
}
