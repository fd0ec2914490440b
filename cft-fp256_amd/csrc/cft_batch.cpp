// cftx_program_run_batch: many instances of one program, of uneven length,
// in one call (include/cft_mi355x.h).
//
// Each lane runs steps[i] trips of the program's segment loop (its first
// top-level REPEAT) as the chain of segments a caller of
// cft_program_run_ex would run: segments of the REPEAT's own count and a
// last one of the remainder, each a full run (fresh registers, scratch
// cleared, the state as the scratch-in block, the scratch-out block as the
// next state).  Two executions of that one definition:
//
//  - persistent: one launch of k_seq_batch.  Lanes are sorted longest first
//    and cut into tiles of `lanes_per_wave`; persistent workgroups take
//    tiles from a counter and run every segment of a tile before taking the
//    next.  Optionally captured once into a HIP graph and replayed.
//  - per-op: the plain form - registers and scratch in HBM, the host walks
//    the instruction stream and launches k_seq_op once per instruction over
//    the lanes still running (a prefix, the lanes being sorted).
#include "seq_prog.hpp"

#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

void seq_drop_graph(cft_program *p)
{
    if (p->gexec)
        (void)hipGraphExecDestroy(p->gexec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    memset(p->gkey, 0, sizeof p->gkey);
}

uint32_t wide_default_budget(const cft_program *prog, int width)
{
    const size_t ipw = 64 / width;
    return (uint32_t)std::min<size_t>(65535, (prog->dev->lds_max / 2) / (ipw * prog->esz));
}

WideDev *wide_at(cft_program *prog, int width, uint32_t max_vregs)
{
    const uint32_t budget = max_vregs ? std::min(max_vregs, wide_default_budget(prog, width))
                                      : wide_default_budget(prog, width);
    // the widest window within budget; each (width, window) schedule is built
    // once and shared by every budget that picks it.  The narrowest window
    // (the fewest registers) is built first: a budget it exceeds is refused
    // without building the others.
    auto at = [&](int window) -> WideDev & {
        window = std::max(window, width);
        const uint64_t key = (uint64_t)width << 32 | (uint32_t)window;
        auto it = prog->wide.find(key);
        if (it == prog->wide.end()) {
            it = prog->wide.emplace(key, WideDev()).first;
            if (!seq_wide_compile_at(prog, width, window, &it->second.wp))
                it->second.failed = true;
        }
        return it->second;
    };
    {
        const WideDev &narrow = at(kWideWindows[kWideWindowCount - 1]);
        if (narrow.failed || narrow.wp.nvregs > budget)
            return nullptr;
    }
    for (int i = 0; i < kWideWindowCount; i++) {
        WideDev &wd = at(kWideWindows[i]);
        if (wd.failed)
            return nullptr;
        if (wd.wp.nvregs > budget)
            continue;
        if (!wd.d_recs) {
            const size_t bytes = wd.wp.recs.size() * sizeof(WideRec);
            if (hipMalloc(&wd.d_recs, bytes) != hipSuccess ||
                hipMemcpy(wd.d_recs, wd.wp.recs.data(), bytes, hipMemcpyHostToDevice) !=
                    hipSuccess) {
                (void)hipGetLastError();
                if (wd.d_recs)
                    (void)hipFree(wd.d_recs);
                wd.d_recs = nullptr;
                return nullptr;
            }
        }
        return &wd;
    }
    return nullptr;
}

bool prog_wide_eligible(cft_program *prog)
{
    if (prog->wide_eligible < 0)
        prog->wide_eligible = seq_wide_eligible(prog) ? 1 : 0;
    return prog->wide_eligible != 0;
}

namespace {

const char *const kWho = "cftx_program_run_batch";

cft_status refuse(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
cft_status refuse(const char *fmt, ...)
{
    char buf[400];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    cft_set_error("%s: %s", kWho, buf);
    return CFT_ERR_INVALID_ARGUMENT;
}

cft_status check(const cft_program *p, const cftx_batch_args *A)
{
    if (!(p->flags & CFT_PROG_FLAG_SCRATCH_IO) || p->n_in == 0 || p->n_in != p->n_out)
        return refuse("the program must carry its state as a scratch block with as many "
                      "slots out as in (it declares %lu in, %lu out)", (unsigned long)p->n_in,
                      (unsigned long)p->n_out);
    if (p->max_deposits)
        return refuse("the program deposits (max_deposits %lu); a batch carries only state",
                      (unsigned long)p->max_deposits);
    if (p->top_repeat == 0xFFFFFFFFu)
        return refuse("the program has no top-level REPEAT to serve as its segment loop");
    if (!A->state || !A->steps)
        return refuse("%s is NULL", !A->state ? "state" : "steps");
    if (A->n > SIZE_MAX / p->n_in / p->esz || A->state_bytes != A->n * p->n_in * p->esz)
        return refuse("state is %zu bytes and %zu lanes of %lu slots are %zu", A->state_bytes,
                      A->n, (unsigned long)p->n_in, A->n * p->n_in * p->esz);
    if (A->n > 0xFFFFFFFFull)
        return refuse("%zu lanes is more than a batch indexes (2^32 - 1)", A->n);
    if (!A->lane_flags && A->lane_flags_bytes)
        return refuse("lane_flags_bytes = %zu with no lane_flags", A->lane_flags_bytes);
    if (A->lane_flags && A->lane_flags_bytes != A->n)
        return refuse("lane_flags_bytes is %zu; a byte a lane is %zu", A->lane_flags_bytes,
                      A->n);
    if (A->mode != CFTX_BATCH_PERSISTENT && A->mode != CFTX_BATCH_PER_OP)
        return refuse("mode %d is not CFTX_BATCH_PERSISTENT or CFTX_BATCH_PER_OP", A->mode);
    if (A->mode == CFTX_BATCH_PER_OP && A->lane_flags)
        return refuse("the per-op form keeps no per-lane flags; pass no lane_flags");
    if (A->lanes_per_wave > 64)
        return refuse("lanes_per_wave %lu is past a wavefront's 64",
                      (unsigned long)A->lanes_per_wave);
    if (A->lanes_per_instance > 64 || (A->lanes_per_instance & (A->lanes_per_instance - 1)))
        return refuse("lanes_per_instance %lu is not a power of two up to 64",
                      (unsigned long)A->lanes_per_instance);
    if (A->waves_per_group != 0 && A->waves_per_group != 1 && A->waves_per_group != 2 &&
        A->waves_per_group != 4)
        return refuse("waves_per_group %lu is not 1, 2 or 4", (unsigned long)A->waves_per_group);
    return CFT_OK;
}

// The per-op form walks only what it can issue one instruction at a time.
cft_status check_per_op(const cft_program *p)
{
    for (uint32_t pc = 0; pc < p->n_insns; pc++) {
        const SeqInsn &in = p->code[pc];
        if (!SEQ_W_CTRL(in.w))
            continue;
        switch (SEQ_W_OP(in.w)) {
        case SEQ_HALT: case SEQ_REPEAT: case SEQ_ENDREP: case SEQ_STL: case SEQ_LDL:
        case SEQ_STX: case SEQ_LDX:
            break;
        default:
            cft_set_error("%s: the per-op form issues arithmetic, loops and scratch "
                          "accesses; instruction %lu is control code %lu", kWho,
                          (unsigned long)pc, (unsigned long)SEQ_W_OP(in.w));
            return CFT_ERR_UNSUPPORTED;
        }
    }
    return CFT_OK;
}

struct PerOp {
    cft_program *p;
    cft_device *dev;
    void *regs, *pool;
    uint64_t stride;
    const uint32_t *bank;
    // trips of the segment loop this round, by position (non-increasing)
    const std::vector<uint32_t> *trips;
    uint64_t launches = 0;

    // positions whose trips exceed k: a prefix
    uint64_t running(uint32_t k) const
    {
        const std::vector<uint32_t> &t = *trips;
        uint64_t lo = 0, hi = t.size();
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if (t[mid] > k) lo = mid + 1; else hi = mid;
        }
        return lo;
    }

    hipError_t walk(uint32_t lo, uint32_t hi, uint64_t n, bool *halted)
    {
        for (uint32_t pc = lo; pc < hi && !*halted; pc++) {
            const SeqInsn &in = p->code[pc];
            if (SEQ_W_CTRL(in.w) && SEQ_W_OP(in.w) == SEQ_HALT) {
                *halted = true;
                break;
            }
            if (SEQ_W_CTRL(in.w) && SEQ_W_OP(in.w) == SEQ_REPEAT) {
                const uint32_t end = in.aux;
                if (pc == p->top_repeat) {
                    const uint32_t most = (*trips)[0];
                    for (uint32_t k = 0; k < most && !*halted; k++) {
                        const uint64_t nk = std::min<uint64_t>(n, running(k));
                        if (nk == 0)
                            break;
                        const hipError_t e = walk(pc + 1, end, nk, halted);
                        if (e != hipSuccess)
                            return e;
                    }
                } else {
                    for (uint32_t k = 0; k < in.imm && !*halted; k++) {
                        const hipError_t e = walk(pc + 1, end, n, halted);
                        if (e != hipSuccess)
                            return e;
                    }
                }
                pc = end;                 // the loop's ENDREP; the for loop steps past it
                continue;
            }
            const hipError_t e =
                cftk_seq_op(p->fmt, &in, regs, pool, stride, n, bank,
                            (p->flags & CFT_PROG_FLAG_SCRATCH_STRICT) ? 1 : 0, p->depth,
                            p->log2depth, dev->dwords, dev->stream);
            launches++;
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
};


// Lanes an instance takes when the caller leaves it open: the width whose
// schedule minimises (steps a segment body) x max(longest lane, total steps
// over the instances the machine holds at once), one wavefront a SIMD.
int choose_width(cft_program *prog, const uint32_t *steps, size_t n)
{
    uint64_t most = 0, total = 0;
    for (size_t i = 0; i < n; i++) {
        most = std::max<uint64_t>(most, steps[i]);
        total += steps[i];
    }
    const double waves = (double)prog->dev->num_cus * 4;
    const uint32_t narrow_body = prog->code[prog->top_repeat].aux - prog->top_repeat - 1;
    double best = (double)narrow_body * std::max<double>((double)most, total / (waves * 64));
    int pick = 1;
    for (int w = 2; w <= 64; w *= 2) {
        WideDev *wd = wide_at(prog, w);
        if (!wd)
            continue;
        const double cap = waves * (64 / w);
        const double cost = 1.1 * wd->wp.body_steps * std::max<double>((double)most, total / cap);
        if (cost < best * 0.97) {
            best = cost;
            pick = w;
        }
    }
    return pick;
}

} // namespace

// The persistent batched launch on `ex`'s stream, device-side inputs bound:
// order/steps uploaded from the host arrays (kept alive by the caller until
// the stream passes), the width chosen or taken, the runner launched (through
// a captured graph when asked).  Does not wait.
cft_status batch_persistent(cft_device *ex, cft_program *prog, const BatchIO &io,
                            const char *who, int *width_out)
{
    const size_t n = io.n, esz = prog->esz;
    const uint32_t seg = prog->code[prog->top_repeat].imm;
    const uint32_t reach = seq_scratch_reach(prog);
    const void *dbank = io.bank, *da = io.a, *db = io.b, *dc = io.c;
    void *dstate = io.state, *dlflags = io.lflags;
    const void *dorder = nullptr, *dsteps = nullptr;
    // the batch takes its longest lanes' time (order is longest first): the
    // waves holding lanes of at least half as many trips issue first
    uint32_t hi_steps = 0;
    if (n > 1 && io.steps && io.order) {
        const uint32_t most = io.steps[io.order[0]];
        hi_steps = most > 1 ? (most + 1) / 2 : 0;
    }
    {
        cft_status bs = bind_in(ex, ROLE_ORDER, io.order, n * 4, &dorder, who);
        if (bs == CFT_OK)
            bs = bind_in(ex, ROLE_TRIPS, io.steps, n * 4, &dsteps, who);
        if (bs != CFT_OK)
            return bs;
    }
    int width = 1;
    {
        width = (int)io.lanes_per_instance;
        if (width > 1 && !prog_wide_eligible(prog)) {
            cft_set_error("%s: lanes_per_instance %d needs a pure program (no deposits, "
                          "activity, indexed scratch or flag control)", who, width);
            return CFT_ERR_UNSUPPORTED;
        }
        if (width == 0)
            width = prog->wide_eligible ? choose_width(prog, io.steps, n) : 1;
    }
    WideDev *wd = width > 1 ? wide_at(prog, width) : nullptr;
    hipError_t e = hipSuccess;
    cft_status st = CFT_OK;
    if (width > 1 && !wd) {
        cft_set_error("%s: the program does not schedule over %d lanes within this "
                      "device's LDS", who, width);
        return CFT_ERR_UNSUPPORTED;
    }
    if (wd) {
        const WideProgram &wp = wd->wp;
        const uint32_t ipw = 64 / (uint32_t)width;
        const bool recs_lds = (size_t)wp.nsteps * width * 16 <= 16384;
        int waves = io.waves_per_group ? (int)io.waves_per_group : 4;
        while (waves > 1 && cftk_seq_wide_lds(prog->fmt, prog->n_consts,
                                              recs_lds ? wp.nsteps : 0, width, wp.nvregs,
                                              waves) > ex->lds_max)
            waves /= 2;
        if (!io.waves_per_group) {
            // the group size that keeps the most waves resident on a CU
            int best = 0;
            for (int w = waves; w >= 1; w /= 2) {
                const int r = w * cftk_seq_wide_blocks_per_cu(prog->fmt, prog->n_consts,
                                                              recs_lds ? wp.nsteps : 0, width,
                                                              wp.nvregs, w);
                if (r > best) {
                    best = r;
                    waves = w;
                }
            }
        }
        if (cftk_seq_wide_lds(prog->fmt, prog->n_consts, recs_lds ? wp.nsteps : 0, width,
                              wp.nvregs, waves) > ex->lds_max) {
            cft_set_error("%s: %lu registers an instance over %d lanes do not fit one "
                          "workgroup's LDS", who, (unsigned long)wp.nvregs, width);
            return CFT_ERR_UNSUPPORTED;
        }
        const int per_cu = cftk_seq_wide_blocks_per_cu(
            prog->fmt, prog->n_consts, recs_lds ? wp.nsteps : 0, width, wp.nvregs, waves);
        const uint64_t ntiles = (n + ipw - 1) / ipw;
        const uint64_t blocks = std::min<uint64_t>((ntiles + waves - 1) / waves,
                                                   (uint64_t)ex->num_cus * per_cu);
        const uint64_t npool = blocks * waves * ipw;
        if (reach && (st = arena_reserve(ex, ex->stage[ROLE_POOL], npool * reach * esz,
                                         who)) != CFT_OK)
            return st;
        SeqArgs S;
        memset(&S, 0, sizeof S);
        S.n_consts = prog->n_consts;
        S.bank = (const uint32_t *)dbank;
        S.a = (const uint32_t *)da;
        S.b = (const uint32_t *)db;
        S.c = (const uint32_t *)dc;
        S.n_in = prog->n_in;
        S.n_out = prog->n_out;
        S.pool = reach ? (uint32_t *)ex->stage[ROLE_POOL].p : nullptr;
        S.npool = npool;
        S.reach = reach;
        S.n = n;
        S.words = ex->dwords;
        SeqBatch B;
        memset(&B, 0, sizeof B);
        B.order = (const uint32_t *)dorder;
        B.steps = (const uint32_t *)dsteps;
        B.state = (uint32_t *)dstate;
        B.queue = ex->dwords + 2;
        B.lflags = (uint8_t *)dlflags;
        B.n = n;
        B.ntiles = (uint32_t)ntiles;
        B.lpw = ipw;
        B.seg = seg;
        B.hi_steps = hi_steps;
        SeqWide Wd;
        memset(&Wd, 0, sizeof Wd);
        Wd.recs = reinterpret_cast<const uint4 *>(wd->d_recs);
        Wd.nsteps = wp.nsteps;
        Wd.width = (uint32_t)width;
        Wd.log2w = 0;
        while ((1u << Wd.log2w) < (uint32_t)width)
            Wd.log2w++;
        Wd.nvregs = wp.nvregs;
        Wd.top_step = wp.top_step;
        Wd.recs_in_lds = recs_lds ? 1 : 0;
        auto record = [&]() -> hipError_t {
            hipError_t r = cftk_fill32(ex->dwords, 0u, 3, ex->stream);
            if (r == hipSuccess)
                r = cftk_seq_wide(prog->fmt, &S, &B, &Wd, prog->phys_abc[0], prog->phys_abc[1],
                                  prog->phys_abc[2], waves, (int)blocks, ex->stream);
            return r;
        };
        e = timer_begin(ex);
        if (e == hipSuccess && io.graph) {
            const uint64_t key[24] = {(uint64_t)(uintptr_t)dorder, (uint64_t)(uintptr_t)dsteps,
                                      (uint64_t)(uintptr_t)dstate, (uint64_t)(uintptr_t)dbank,
                                      (uint64_t)(uintptr_t)da, (uint64_t)(uintptr_t)db,
                                      (uint64_t)(uintptr_t)dc, (uint64_t)(uintptr_t)dlflags,
                                      (uint64_t)(uintptr_t)S.pool, n, ipw, (uint64_t)waves,
                                      blocks, npool, seg, reach, 2, (uint64_t)width,
                                      (uint64_t)(uintptr_t)wd->d_recs, hi_steps, 0, 0, 0, 0};
            if (!prog->gexec || memcmp(key, prog->gkey, sizeof key) != 0) {
                seq_drop_graph(prog);
                e = hipStreamBeginCapture(ex->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    const hipError_t r = record();
                    e = hipStreamEndCapture(ex->stream, &prog->graph);
                    if (r != hipSuccess)
                        e = r;
                }
                if (e == hipSuccess)
                    e = hipGraphInstantiate(&prog->gexec, prog->graph, nullptr, nullptr, 0);
                if (e == hipSuccess)
                    memcpy(prog->gkey, key, sizeof key);
            }
            if (e == hipSuccess)
                e = hipGraphLaunch(prog->gexec, ex->stream);
        } else if (e == hipSuccess) {
            e = record();
        }
        if (e == hipSuccess)
            e = timer_end(ex);
        if (e != hipSuccess)
            return dev_fail(ex, e, who, "launching the wide batched runner");
    } else {
        const Shape sh = shape_for(prog, io.waves_per_group ? (int)io.waves_per_group : 4);
        if (sh.waves == 0) {
            cft_set_error("%s: this program's %lu registers do not fit one workgroup's LDS",
                          who, (unsigned long)prog->nregs);
            return CFT_ERR_UNSUPPORTED;
        }
        const int per_cu = cftk_seq_batch_blocks_per_cu(
            prog->fmt, sh.code_in_lds ? prog->n_insns : 0, prog->n_consts, prog->nregs,
            sh.waves);
        const uint64_t slots = (uint64_t)ex->num_cus * per_cu * sh.waves;
        uint32_t lpw = io.lanes_per_wave;
        if (lpw == 0)      // fill every wave slot before packing lanes densely
            lpw = (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(1, (n + slots - 1) / slots));
        const uint64_t ntiles = (n + lpw - 1) / lpw;
        const uint64_t blocks =
            std::min<uint64_t>((ntiles + sh.waves - 1) / sh.waves, (uint64_t)ex->num_cus * per_cu);
        const uint64_t npool = blocks * sh.waves * 64;
        if (reach && (st = arena_reserve(ex, ex->stage[ROLE_POOL], npool * reach * esz,
                                         who)) != CFT_OK)
            return st;
        SeqArgs S;
        memset(&S, 0, sizeof S);
        S.code = prog->d_code;
        S.n_insns = prog->n_insns;
        S.n_consts = prog->n_consts;
        S.bank = (const uint32_t *)dbank;
        S.a = (const uint32_t *)da;
        S.b = (const uint32_t *)db;
        S.c = (const uint32_t *)dc;
        S.n_in = prog->n_in;
        S.n_out = prog->n_out;
        S.pool = reach ? (uint32_t *)ex->stage[ROLE_POOL].p : nullptr;
        S.npool = npool;
        S.reach = reach;
        S.depth = prog->depth;
        S.log2depth = prog->log2depth;
        S.strict = (prog->flags & CFT_PROG_FLAG_SCRATCH_STRICT) ? 1 : 0;
        S.n = n;
        S.words = ex->dwords;
        S.code_in_lds = sh.code_in_lds ? 1 : 0;
        SeqBatch B;
        memset(&B, 0, sizeof B);
        B.order = (const uint32_t *)dorder;
        B.steps = (const uint32_t *)dsteps;
        B.state = (uint32_t *)dstate;
        B.queue = ex->dwords + 2;
        B.lflags = (uint8_t *)dlflags;
        B.n = n;
        B.ntiles = (uint32_t)ntiles;
        B.lpw = lpw;
        B.seg = seg;
        B.hi_steps = hi_steps;
        auto record = [&]() -> hipError_t {
            hipError_t r = cftk_fill32(ex->dwords, 0u, 3, ex->stream);
            if (r == hipSuccess)
                r = cftk_seq_batch(prog->fmt, &S, &B, prog->nregs, prog->phys_abc[0],
                                   prog->phys_abc[1], prog->phys_abc[2], prog->top_repeat,
                                   sh.waves, (int)blocks, ex->stream);
            return r;
        };
        e = timer_begin(ex);
        if (e == hipSuccess && io.graph) {
            const uint64_t key[24] = {(uint64_t)(uintptr_t)dorder, (uint64_t)(uintptr_t)dsteps,
                                      (uint64_t)(uintptr_t)dstate, (uint64_t)(uintptr_t)dbank,
                                      (uint64_t)(uintptr_t)da, (uint64_t)(uintptr_t)db,
                                      (uint64_t)(uintptr_t)dc, (uint64_t)(uintptr_t)dlflags,
                                      (uint64_t)(uintptr_t)S.pool, n, lpw, (uint64_t)sh.waves, blocks,
                                      npool, seg, reach, 1, hi_steps, 0, 0, 0, 0, 0, 0};
            if (!prog->gexec || memcmp(key, prog->gkey, sizeof key) != 0) {
                seq_drop_graph(prog);
                e = hipStreamBeginCapture(ex->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    const hipError_t r = record();
                    e = hipStreamEndCapture(ex->stream, &prog->graph);
                    if (r != hipSuccess)
                        e = r;
                }
                if (e == hipSuccess)
                    e = hipGraphInstantiate(&prog->gexec, prog->graph, nullptr, nullptr, 0);
                if (e == hipSuccess)
                    memcpy(prog->gkey, key, sizeof key);
            }
            if (e == hipSuccess)
                e = hipGraphLaunch(prog->gexec, ex->stream);
        } else if (e == hipSuccess) {
            e = record();
        }
        if (e == hipSuccess)
            e = timer_end(ex);
        if (e != hipSuccess)
            return dev_fail(ex, e, who, "launching the batched runner");
    }
    *width_out = width;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_program_run_batch(cft_program *prog,
                                                     const cftx_batch_args *args)
{
    cft_clear_error();
    if (!prog || !args) {
        cft_set_error("%s: %s is NULL", kWho, !prog ? "prog" : "args");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (args->struct_size != sizeof(cftx_batch_args)) {
        cft_set_error("%s was given a %zu-byte cftx_batch_args and this library's is %zu", kWho,
                      args->struct_size, sizeof(cftx_batch_args));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const cftx_batch_args &A = *args;
    if (A.bus_out)
        *A.bus_out = 0;
    cft_device *dev = prog->dev;
    if (A.n == 0) {
        cft_emit_flags(dev, 0, A.flags_out);
        return CFT_OK;
    }
    cft_status st = check(prog, &A);
    if (st == CFT_OK)
        st = seq_check_bank(prog, A.bank, A.bank_bytes, kWho);
    if (st == CFT_OK && A.mode == CFTX_BATCH_PER_OP)
        st = check_per_op(prog);
    if (st == CFT_OK)
        st = host_in(dev, A.steps, kWho);
    if (st != CFT_OK)
        return st;
    st = dev_enter(dev, kWho);
    if (st != CFT_OK)
        return st;

    const size_t n = A.n, esz = prog->esz;
    const uint32_t seg = prog->code[prog->top_repeat].imm;
    const uint32_t reach = seq_scratch_reach(prog);

    // longest first
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return A.steps[x] > A.steps[y]; });

    const void *dorder = nullptr, *dsteps = nullptr, *dbank = nullptr, *da = nullptr,
               *db = nullptr, *dc = nullptr;
    if ((A.mode == CFTX_BATCH_PER_OP &&
         ((st = bind_in(dev, ROLE_ORDER, order.data(), n * 4, &dorder, kWho)) != CFT_OK ||
          (st = bind_in(dev, ROLE_TRIPS, A.steps, n * 4, &dsteps, kWho)) != CFT_OK)) ||
        (st = bind_in(dev, ROLE_A, A.a, n * esz, &da, kWho)) != CFT_OK ||
        (st = bind_in(dev, ROLE_B, A.b, n * esz, &db, kWho)) != CFT_OK ||
        (st = bind_in(dev, ROLE_C, A.c, n * esz, &dc, kWho)) != CFT_OK)
        return st;
    if (prog->flags & CFT_PROG_FLAG_BANK_EXT) {
        if (A.bank_bytes &&
            (st = bind_in(dev, ROLE_BANK, A.bank, A.bank_bytes, &dbank, kWho)) != CFT_OK)
            return st;
    } else {
        dbank = prog->d_consts;
    }
    // the state: read and written in place
    void *dstate = nullptr;
    cft_buffer *sbuf = nullptr;
    if ((st = bind_out(dev, ROLE_SIN, A.state, A.state_bytes, &dstate, &sbuf, kWho)) != CFT_OK)
        return st;
    hipError_t e = hipSuccess;
    if (!sbuf)
        e = hipMemcpyAsync(dstate, A.state, A.state_bytes, hipMemcpyHostToDevice, dev->stream);
    void *dlflags = nullptr;
    cft_buffer *lbuf = nullptr;
    if (e == hipSuccess && A.lane_flags &&
        (st = bind_out(dev, ROLE_LFLAGS, A.lane_flags, n, &dlflags, &lbuf, kWho)) != CFT_OK)
        return st;
    if (e != hipSuccess)
        return dev_fail(dev, e, kWho, "staging the state");

    uint64_t launches = 0;
    int width = 1;
    if (A.mode == CFTX_BATCH_PERSISTENT) {
        BatchIO io;
        io.bank = dbank;
        io.a = da;
        io.b = db;
        io.c = dc;
        io.state = dstate;
        io.lflags = (uint8_t *)dlflags;
        io.steps = A.steps;
        io.order = order.data();
        io.n = n;
        io.lanes_per_wave = A.lanes_per_wave;
        io.waves_per_group = A.waves_per_group;
        io.lanes_per_instance = A.lanes_per_instance;
        io.graph = A.graph != 0;
        st = batch_persistent(dev, prog, io, kWho, &width);
        if (st != CFT_OK)
            return st;
        launches = 2;
    } else {
        // the plain form
        const uint64_t stride = n;
        if ((st = arena_reserve(dev, dev->stage[ROLE_REGS], prog->nregs * n * esz, kWho)) !=
                CFT_OK ||
            (reach && (st = arena_reserve(dev, dev->stage[ROLE_POOL], reach * n * esz, kWho)) !=
                          CFT_OK))
            return st;
        PerOp W;
        W.p = prog;
        W.dev = dev;
        W.regs = dev->stage[ROLE_REGS].p;
        W.pool = reach ? dev->stage[ROLE_POOL].p : nullptr;
        W.stride = stride;
        W.bank = (const uint32_t *)dbank;
        std::vector<uint32_t> left(n), trips(n);
        for (size_t i = 0; i < n; i++)
            left[i] = A.steps[order[i]];
        W.trips = &trips;
        e = cftk_fill32(dev->dwords, 0u, 3, dev->stream);
        if (e == hipSuccess)
            e = timer_begin(dev);
        while (e == hipSuccess) {
            uint64_t nr = 0;
            while (nr < n && left[nr] > 0)
                nr++;
            if (nr == 0)
                break;
            trips.assign(left.begin(), left.begin() + nr);
            for (uint64_t i = 0; i < nr; i++)
                trips[i] = std::min(trips[i], seg);
            e = cftk_seq_lanes(prog->fmt, W.regs, W.pool, stride, nr, prog->nregs, reach,
                               prog->n_in, (const uint32_t *)dorder, (uint32_t *)dstate,
                               (const uint32_t *)da, (const uint32_t *)db, (const uint32_t *)dc,
                               prog->phys_abc[0], prog->phys_abc[1], prog->phys_abc[2], 0,
                               dev->stream);
            bool halted = false;
            if (e == hipSuccess)
                e = W.walk(0, prog->n_insns, nr, &halted);
            if (e == hipSuccess)
                e = cftk_seq_lanes(prog->fmt, W.regs, W.pool, stride, nr, prog->nregs, reach,
                                   prog->n_out, (const uint32_t *)dorder, (uint32_t *)dstate,
                                   nullptr, nullptr, nullptr, -1, -1, -1, 1, dev->stream);
            launches += 2;
            for (uint64_t i = 0; i < nr; i++)
                left[i] -= trips[i];
        }
        launches += W.launches;
        if (e == hipSuccess)
            e = timer_end(dev);
        if (e != hipSuccess)
            return dev_fail(dev, e, kWho, "launching the per-op form");
    }
    if (!sbuf)
        e = hipMemcpyAsync(A.state, dstate, A.state_bytes, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess && dlflags && !lbuf)
        e = hipMemcpyAsync(A.lane_flags, dlflags, n, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dev->hwords, dev->dwords, 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, kWho, "running the batch");
    timer_collect(dev);
    if (sbuf)
        sbuf->host_valid = false;
    if (lbuf)
        lbuf->host_valid = false;
    dev->batch_launches = launches;
    dev->batch_width = A.mode == CFTX_BATCH_PERSISTENT ? (uint32_t)width : 0;
    cft_emit_flags(dev, dev->hwords[0], A.flags_out);
    if (A.bus_out)
        *A.bus_out = dev->hwords[1];
    return CFT_OK;
}

extern "C" CFT_API uint64_t cftx_batch_launches(cft_device *dev)
{
    return dev ? dev->batch_launches : 0;
}

extern "C" CFT_API uint32_t cftx_batch_width(cft_device *dev)
{
    return dev ? dev->batch_width : 0;
}
