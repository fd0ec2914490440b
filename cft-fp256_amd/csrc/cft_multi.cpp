// Instances of several sequencer programs in one launch: the planner and
// launcher of k_seq_multi (multi_launch, used by cftx_program_run_multi and
// the serving runtime), cftx_program_run_multi and cftx_program_prepare
// (include/cft_mi355x.h).
//
// Every lane runs its trips of its program's segment loop as the chain
// cftx_program_run_batch defines (segments of the REPEAT's own count and a
// last one of the remainder), over the program's wide schedule at the width
// the planner gives that program.  A lane's results depend on its own state
// and trips only, so they are those of the same lane run alone.
//
// The planner.  For a program p at width W (64 / W instances a wave), its
// lanes sorted longest first and cut into tiles of 64 / W, a tile costs
// (its longest lane's trips) x (bundles a trip at W) in bundles of one wave.
// The launch's makespan is taken as
//     max(sum of tile costs / issue slots, longest tile / a)
// where a wave alone on a SIMD issues at a fraction a (0.9) of the SIMD's
// rate and two or more keep it busy, so a CU gives sum over its 4 SIMDs of
// min(1, a x waves on it).  Waves a CU holds follow from the LDS region a
// wave needs, the largest over the programs of
//     (constants + nvregs(W) x 64 / W) rows,
// and a program's candidates at each W include schedules built within
// register budgets that let 4, 6 or 8 waves share a CU's LDS.
// The planner tries each cap on the region (the waves a CU must hold, 1 to
// 12) and each distinct longest-tile bound L: every program takes its
// cheapest candidate within the cap whose longest tile is within L (the
// fastest within the cap if none is), and the plan with the smallest
// makespan wins.  The caller may fix the group size, the waves a CU holds
// and the lone-wave fraction a: cftx_program_run_multi's options, which an
// autotuner sweeps on a sample.
#include "seq_prog.hpp"

#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

struct Cand {
    uint32_t width = 0;
    double work = 0, longest = 0;
    uint32_t rows = 0;
    WideDev *wd = nullptr;
};

// Rows of the format's LDS row type an element takes.
uint32_t rows_per_el(int fmt)
{
    return fmt == 3 ? 2u : 1u;
}

size_t wave_lds(int fmt, uint32_t rows)
{
    return (size_t)rows * cftk_seq_row_bytes(fmt) + 4 * 64 + 8;
}

// A wave alone on its SIMD issues at about this fraction of the SIMD's rate,
// a format: a binary256 bundle is long straight-line integer code, so one
// wave nearly fills the SIMD (measured: one Lorenz-63 wave alone takes 2.55
// us a bundle, a full ensemble 2.86 us a bundle a SIMD); the short bundles of
// the narrow formats leave LDS and record-load latency exposed.
constexpr double kAlone[4] = {0.4, 0.5, 0.75, 0.9};

// SIMDs' worth of full-rate issue the device gives when every wave needs
// `rows` rows of LDS: waves a CU holds (LDS, and 12 for registers), spread
// over its 4 SIMDs, each SIMD at min(1, alone x its waves).
double issue_slots(const cft_device *ex, int fmt, uint32_t rows, double lds_cu, double alone)
{
    const int per_cu = (int)std::min(12.0, std::floor(lds_cu / (double)wave_lds(fmt, rows)));
    double cu = 0;
    for (int s = 0; s < 4; s++) {
        const int k = per_cu / 4 + (s < per_cu % 4 ? 1 : 0);
        cu += std::min(1.0, alone * k);
    }
    return (double)ex->num_cus * std::max(cu, alone);
}

// Register budgets a program is scheduled within at width w: the default,
// and tighter ones that let 4, 6 or 8 waves share a CU's LDS (fewer
// registers, usually a few more bundles).
std::vector<uint32_t> budgets(const cft_program *p, int w)
{
    std::vector<uint32_t> out{0};
    const uint32_t ipw = 64 / (uint32_t)w, nq = rows_per_el(p->fmt);
    for (int per_cu : {4, 6, 8}) {
        const double rows = ((double)p->dev->lds_max / per_cu - 4 * 64 - 8) /
                            (double)cftk_seq_row_bytes(p->fmt);
        const double v = (rows / nq - p->n_consts) / ipw;
        if (v >= 1 && v < wide_default_budget(p, w))
            out.push_back((uint32_t)v);
    }
    return out;
}

} // namespace

PinnedStage::~PinnedStage()
{
    if (p)
        (void)hipHostFree(p);
    for (uint8_t *q : retired)
        (void)hipHostFree(q);
}

cft_status multi_launch(cft_device *ex, std::vector<MultiPart> &parts, uint8_t *dlflags,
                        const MultiTune &tune, MultiRun &run, const char *who,
                        PinnedStage *pin)
{
    const uint32_t waves_per_group = tune.waves_per_group;
    if (parts.empty())
        return CFT_OK;
    const int fmt = parts[0].prog->fmt;
    const size_t esz = parts[0].prog->esz;
    const uint32_t nq = rows_per_el(fmt);
    const size_t np = parts.size();

    // each part's lanes with trips, longest first
    std::vector<std::vector<uint32_t>> lanes(np);
    for (size_t p = 0; p < np; p++) {
        MultiPart &P = parts[p];
        std::vector<uint32_t> &L = lanes[p];
        for (uint32_t i = 0; i < (uint32_t)P.steps.size(); i++)
            if (P.steps[i])
                L.push_back(i);
        std::stable_sort(L.begin(), L.end(),
                         [&](uint32_t x, uint32_t y) { return P.steps[x] > P.steps[y]; });
    }

    run = MultiRun();
    {
        bool any = false;
        for (const std::vector<uint32_t> &L : lanes)
            any |= !L.empty();
        if (!any)
            return CFT_OK;             // no lane has trips: nothing to launch
    }

    // candidates a part
    std::vector<std::vector<Cand>> cands(np);
    for (size_t p = 0; p < np; p++) {
        MultiPart &P = parts[p];
        const std::vector<uint32_t> &L = lanes[p];
        if (L.empty())
            continue;
        uint32_t fewest = 0xFFFFFFFFu;          // the fewest registers a schedule took
        for (uint32_t w = 64; w >= 1; w /= 2) {
            if (P.width && P.width != w)
                continue;
            const uint32_t ipw = 64 / w;
            if (fewest != 0xFFFFFFFFu &&
                wave_lds(fmt, (P.prog->n_consts + fewest * ipw) * nq) > ex->lds_max)
                continue;                      // narrower widths hold more instances a wave
            double trips = 0;
            for (size_t k = 0; k < L.size(); k += ipw)
                trips += P.steps[L[k]];
            std::vector<uint32_t> seen;
            for (uint32_t b : budgets(P.prog, (int)w)) {
                WideDev *wd = wide_at(P.prog, (int)w, b);
                if (!wd || std::find(seen.begin(), seen.end(), wd->wp.nvregs) != seen.end())
                    continue;
                seen.push_back(wd->wp.nvregs);
                fewest = std::min(fewest, wd->wp.nvregs);
                Cand c;
                c.width = w;
                c.wd = wd;
                const double body = std::max<uint32_t>(1, wd->wp.body_steps);
                c.work = trips * body;
                c.longest = (double)P.steps[L[0]] * body;
                c.rows = (P.prog->n_consts + wd->wp.nvregs * ipw) * nq;
                if (wave_lds(fmt, c.rows) > ex->lds_max)
                    continue;
                cands[p].push_back(c);
            }
        }
        if (cands[p].empty()) {
            if (P.width)
                cft_set_error("%s: program %zu does not schedule over %lu lanes within one "
                              "workgroup's LDS", who, p, (unsigned long)P.width);
            else
                cft_set_error("%s: program %zu schedules at no width within one workgroup's "
                              "LDS", who, p);
            return CFT_ERR_UNSUPPORTED;
        }
    }

    // the search: for each cap on the LDS region (the waves a CU must hold)
    // and each longest-tile bound, every program takes its cheapest candidate
    // within both (the fastest within the cap if none is within the bound)
    const double alone = tune.alone > 0 ? std::min(1.0, tune.alone) : kAlone[fmt];
    std::vector<double> bounds;
    for (auto &cs : cands)
        for (const Cand &c : cs)
            bounds.push_back(c.longest);
    std::sort(bounds.begin(), bounds.end());
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    const double lds_cu = (double)ex->lds_max;
    std::vector<uint32_t> caps;
    for (int k = 1; k <= 12; k++) {
        if (tune.min_waves_per_cu && k != (int)tune.min_waves_per_cu)
            continue;
        const double r = (lds_cu / k - 4 * 64 - 8) / (double)cftk_seq_row_bytes(fmt);
        if (r >= 1)
            caps.push_back((uint32_t)r);
    }
    // each part's candidates within a cap, by longest tile, with the running
    // cheapest: the cheapest within a bound is one binary search away
    struct Ranked {
        std::vector<double> longest;    // ascending
        std::vector<int> idx;           // candidate, same order
        std::vector<int> cheapest;      // cheapest among [0, i]
    };
    std::vector<int> best_pick(np, -1);
    double best = -1, best_long = 0;
    uint32_t best_rows = 0;
    for (uint32_t cap : caps) {
        std::vector<Ranked> rk(np);
        bool fits = true;
        for (size_t p = 0; p < np && fits; p++) {
            const std::vector<Cand> &cs = cands[p];
            if (cs.empty())
                continue;
            Ranked &R = rk[p];
            for (int i = 0; i < (int)cs.size(); i++)
                if (cs[i].rows <= cap)
                    R.idx.push_back(i);
            if (R.idx.empty()) {
                fits = false;
                break;
            }
            std::sort(R.idx.begin(), R.idx.end(),
                      [&](int a, int b) { return cs[a].longest < cs[b].longest; });
            int c = -1;
            for (int i : R.idx) {
                R.longest.push_back(cs[i].longest);
                if (c < 0 || cs[i].work < cs[c].work ||
                    (cs[i].work == cs[c].work && cs[i].rows < cs[c].rows))
                    c = i;
                R.cheapest.push_back(c);
            }
        }
        if (!fits)
            continue;
        for (double bound : bounds) {
            double work = 0, longest = 0;
            uint32_t rows = 0;
            std::vector<int> pick(np, -1);
            for (size_t p = 0; p < np; p++) {
                const Ranked &R = rk[p];
                if (R.idx.empty())
                    continue;
                const size_t n_in =
                    std::upper_bound(R.longest.begin(), R.longest.end(), bound) -
                    R.longest.begin();
                // the cheapest within the bound, else the fastest within the cap
                const int k = n_in ? R.cheapest[n_in - 1] : R.idx[0];
                const Cand &c = cands[p][k];
                pick[p] = k;
                work += c.work;
                longest = std::max(longest, c.longest);
                rows = std::max(rows, c.rows);
            }
            const double span =
                std::max(work / issue_slots(ex, fmt, rows, lds_cu, alone), longest / alone);
            if (best < 0 || span < best * 0.999 || (span <= best * 1.001 && rows < best_rows)) {
                best = span;
                best_long = longest;
                best_rows = rows;
                best_pick = pick;
            }
        }
    }
    if (best < 0) {
        cft_set_error("%s: no plan holds %lu waves a CU (min_waves_per_cu)", who,
                      (unsigned long)tune.min_waves_per_cu);
        return CFT_ERR_UNSUPPORTED;
    }

    // positions, tiles and program entries
    run = MultiRun();
    run.body.assign(np, 0.0);
    run.makespan = best;
    run.longest = best_long;
    run.region_rows = best_rows;
    uint32_t reach_max = 0;
    struct TileC {
        uint32_t prog, first, count;
        double cost;
    };
    std::vector<TileC> tl;
    for (size_t p = 0; p < np; p++) {
        MultiPart &P = parts[p];
        cft_program *prog = P.prog;
        const std::vector<uint32_t> &L = lanes[p];
        if (best_pick[p] < 0) {
            if (!P.width)
                P.width = 1;
            continue;
        }
        const Cand &c = cands[p][best_pick[p]];
        P.width = c.width;
        const WideProgram &wp = c.wd->wp;
        const uint32_t pi = (uint32_t)run.progs.size();
        SeqMProg mp;
        memset(&mp, 0, sizeof mp);
        mp.recs = reinterpret_cast<const uint4 *>(c.wd->d_recs);
        mp.bank = (const uint32_t *)P.dbank;
        mp.nsteps = wp.nsteps;
        mp.log2w = 0;
        while ((1u << mp.log2w) < c.width)
            mp.log2w++;
        mp.nvregs = wp.nvregs;
        mp.top_step = wp.top_step;
        mp.n_consts = prog->n_consts;
        mp.n_in = prog->n_in;
        mp.reach = seq_scratch_reach(prog);
        mp.seg = prog->code[prog->top_repeat].imm;
        reach_max = std::max(reach_max, mp.reach);
        run.progs.push_back(mp);
        const uint32_t ipw = 64 / c.width;
        const double body = std::max<uint32_t>(1, wp.body_steps);
        run.body[p] = body;
        const uint32_t base = (uint32_t)run.saddr.size();
        for (uint32_t i : L) {
            run.saddr.push_back(P.saddr[i]);
            run.steps.push_back(P.steps[i]);
            run.part_of.push_back((uint32_t)p);
            run.lane_of.push_back(i);
        }
        for (size_t k = 0; k < L.size(); k += ipw) {
            TileC t;
            t.prog = pi;
            t.first = base + (uint32_t)k;
            t.count = (uint32_t)std::min<size_t>(ipw, L.size() - k);
            t.cost = (double)P.steps[L[k]] * body;
            tl.push_back(t);
        }
    }
    if (tl.empty())
        return CFT_OK;
    std::stable_sort(tl.begin(), tl.end(),
                     [](const TileC &a, const TileC &b) { return a.cost > b.cost; });
    run.tiles.resize(tl.size() * 4);
    for (size_t t = 0; t < tl.size(); t++) {
        run.tiles[4 * t] = tl[t].prog;
        run.tiles[4 * t + 1] = tl[t].first;
        run.tiles[4 * t + 2] = tl[t].count;
        run.tiles[4 * t + 3] = 0;
    }

    // the grid: the largest workgroup (waves) whose LDS fits, persistent blocks
    int waves = waves_per_group ? (int)waves_per_group : 4;
    while (waves > 1 && cftk_seq_multi_lds(fmt, run.region_rows, waves) > ex->lds_max)
        waves /= 2;
    const uint64_t ntiles = tl.size();
    while (waves > 1 && (uint64_t)waves > ntiles)
        waves /= 2;
    if (!waves_per_group) {
        // the group size that keeps the most waves resident on a CU
        int best = 0;
        for (int w = waves; w >= 1; w /= 2) {
            const int r = w * cftk_seq_multi_blocks_per_cu(fmt, run.region_rows, w);
            if (r > best) {
                best = r;
                waves = w;
            }
        }
    }
    int per_cu = cftk_seq_multi_blocks_per_cu(fmt, run.region_rows, waves);
    if (tune.max_waves_per_cu)
        per_cu = std::max(1, std::min(per_cu, (int)tune.max_waves_per_cu / waves));
    run.waves = (uint32_t)waves;
    run.blocks = std::min<uint64_t>((ntiles + waves - 1) / waves, (uint64_t)ex->num_cus * per_cu);
    ex->multi_plan[0] = (uint64_t)waves;
    ex->multi_plan[1] = run.blocks;
    ex->multi_plan[2] = run.region_rows;
    ex->multi_plan[3] = ntiles;
    ex->multi_plan[4] = (uint64_t)per_cu;
    ex->multi_plan[5] = (uint64_t)run.makespan;
    const uint64_t npool = run.blocks * (uint64_t)waves * 64;

    cft_status st = CFT_OK;
    auto reserve = [&](DevArena &a, size_t bytes) {
        return pin ? arena_grow(ex, a, bytes, who) : arena_reserve(ex, a, bytes, who);
    };
    if (reach_max && (st = reserve(ex->stage[ROLE_POOL], npool * reach_max * esz)) != CFT_OK)
        return st;
    // the four tables as one block: programs | tiles | state addresses | trips
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_progs = run.progs.size() * sizeof(SeqMProg), b_tiles = run.tiles.size() * 4,
                 b_saddr = run.saddr.size() * 8, b_steps = run.steps.size() * 4;
    const size_t o_tiles = al(b_progs), o_saddr = o_tiles + al(b_tiles),
                 o_steps = o_saddr + al(b_saddr), total = o_steps + al(b_steps);
    uint8_t *host = nullptr;
    if (pin) {
        if (pin->cap < total) {
            // grow twofold; the old buffer is kept, not freed (a free would
            // wait for the device)
            const size_t want = std::max<size_t>({total, pin->cap * 2, (size_t)1 << 16});
            uint8_t *q = nullptr;
            if (hipHostMalloc((void **)&q, want) != hipSuccess) {
                (void)hipGetLastError();
                cft_set_error("%s: %zu bytes of pinned host memory could not be allocated",
                              who, want);
                return CFT_ERR_OUT_OF_MEMORY;
            }
            if (pin->p)
                pin->retired.push_back(pin->p);
            pin->p = q;
            pin->cap = want;
        }
        host = pin->p;
    } else {
        run.blob.resize(total);
        host = run.blob.data();
    }
    memcpy(host, run.progs.data(), b_progs);
    memcpy(host + o_tiles, run.tiles.data(), b_tiles);
    memcpy(host + o_saddr, run.saddr.data(), b_saddr);
    memcpy(host + o_steps, run.steps.data(), b_steps);
    if ((st = reserve(ex->stage[ROLE_MPROGS], total)) != CFT_OK)
        return st;
    uint8_t *dev_tables = (uint8_t *)ex->stage[ROLE_MPROGS].p;
    hipError_t ce = hipMemcpyAsync(dev_tables, host, total, hipMemcpyHostToDevice, ex->stream);
    if (ce != hipSuccess)
        return dev_fail(ex, ce, who, "copying the launch's tables");
    const void *dprogs = dev_tables, *dtiles = dev_tables + o_tiles,
               *dsaddr = dev_tables + o_saddr, *dsteps = dev_tables + o_steps;
    SeqArgs A;
    memset(&A, 0, sizeof A);
    A.pool = reach_max ? (uint32_t *)ex->stage[ROLE_POOL].p : nullptr;
    A.npool = npool;
    A.n = run.saddr.size();
    A.words = ex->dwords;
    SeqMulti M;
    memset(&M, 0, sizeof M);
    M.progs = (const SeqMProg *)dprogs;
    M.tiles = (const uint4 *)dtiles;
    M.saddr = (const uint64_t *)dsaddr;
    M.steps = (const uint32_t *)dsteps;
    M.lflags = dlflags;
    M.queue = ex->dwords + 2;
    M.ntiles = (uint32_t)ntiles;
    M.region_rows = run.region_rows;
    hipError_t e = cftk_fill32(ex->dwords, 0u, 3, ex->stream);
    if (e == hipSuccess)
        e = timer_begin(ex);
    if (e == hipSuccess)
        e = cftk_seq_multi(fmt, &A, &M, waves, (int)run.blocks, ex->stream);
    if (e == hipSuccess)
        e = timer_end(ex);
    if (e != hipSuccess)
        return dev_fail(ex, e, who, "launching the multi-program runner");
    run.launched = true;
    return CFT_OK;
}

// ---- cftx_program_run_multi -------------------------------------------------

namespace {

const char *const kWho = "cftx_program_run_multi";

cft_status check_part(cft_device *dev, const cftx_batch_part &P, size_t k, int fmt)
{
    const cft_program *p = P.prog;
    auto bad = [&](const char *what) {
        cft_set_error("%s: part %zu: %s", kWho, k, what);
        return CFT_ERR_INVALID_ARGUMENT;
    };
    if (P.struct_size != sizeof(cftx_batch_part))
        return bad("struct_size is not this library's sizeof(cftx_batch_part)");
    if (!p)
        return bad("prog is NULL");
    if (p->dev != dev)
        return bad("its program was loaded on another device handle");
    if (p->fmt != fmt)
        return bad("its program's format differs from part 0's; one launch runs one format");
    if (!(p->flags & CFT_PROG_FLAG_SCRATCH_IO) || p->n_in == 0 || p->n_in != p->n_out)
        return bad("its program must carry its state as a scratch block with as many slots "
                   "out as in");
    if (p->max_deposits)
        return bad("its program deposits; a batch carries only state");
    if (p->top_repeat == 0xFFFFFFFFu)
        return bad("its program has no top-level REPEAT to serve as its segment loop");
    if (P.n && (!P.state || !P.steps))
        return bad("state or steps is NULL");
    if (P.n > 0xFFFFFFFFull || P.n > SIZE_MAX / p->n_in / p->esz ||
        P.state_bytes != P.n * p->n_in * p->esz)
        return bad("state_bytes is not n x slots x the element size");
    if (P.lanes_per_instance > 64 || (P.lanes_per_instance & (P.lanes_per_instance - 1)))
        return bad("lanes_per_instance is not 0 or a power of two up to 64");
    return CFT_OK;
}

} // namespace

extern "C" CFT_API cft_status cftx_program_run_multi(cft_device *dev, cftx_batch_part *parts,
                                                     size_t nparts, const cftx_multi_opts *opts,
                                                     uint32_t *flags_out, uint32_t *bus_out)
{
    cft_clear_error();
    if (bus_out)
        *bus_out = 0;
    if (!dev || !parts || nparts == 0) {
        cft_set_error("%s: %s", kWho, !dev ? "dev is NULL" : "no parts");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    MultiTune tune;
    if (opts) {
        if (opts->struct_size != sizeof(cftx_multi_opts)) {
            cft_set_error("%s was given a %zu-byte cftx_multi_opts and this library's is %zu",
                          kWho, opts->struct_size, sizeof(cftx_multi_opts));
            return CFT_ERR_INVALID_ARGUMENT;
        }
        if ((opts->waves_per_group != 0 && opts->waves_per_group != 1 &&
             opts->waves_per_group != 2 && opts->waves_per_group != 4) ||
            opts->waves_per_cu > 12 || !(opts->alone >= 0 && opts->alone <= 1)) {
            cft_set_error("%s: options out of range (waves_per_group 0, 1, 2 or 4; "
                          "waves_per_cu 0..12; alone 0..1)", kWho);
            return CFT_ERR_INVALID_ARGUMENT;
        }
        tune.waves_per_group = opts->waves_per_group;
        tune.min_waves_per_cu = opts->waves_per_cu;
        tune.alone = opts->alone;
    }
    if (!parts[0].prog || parts[0].struct_size != sizeof(cftx_batch_part)) {
        cft_set_error("%s: part 0 has no program or a foreign struct_size", kWho);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    const int fmt = parts[0].prog->fmt;
    size_t total = 0;
    cft_status st = CFT_OK;
    for (size_t k = 0; k < nparts && st == CFT_OK; k++) {
        st = check_part(dev, parts[k], k, fmt);
        if (st == CFT_OK)
            st = seq_check_bank(parts[k].prog, parts[k].bank, parts[k].bank_bytes, kWho);
        if (st == CFT_OK && parts[k].n)
            st = host_in(dev, parts[k].steps, kWho);
        total += parts[k].n;
    }
    if (st != CFT_OK)
        return st;
    if (total > 0xFFFFFFFFull) {
        cft_set_error("%s: %zu lanes in all is more than a launch indexes", kWho, total);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    for (size_t k = 0; k < nparts; k++)
        parts[k].width_used = 0;
    if (total == 0) {
        cft_emit_flags(dev, 0, flags_out);
        return CFT_OK;
    }
    if ((st = dev_enter(dev, kWho)) != CFT_OK)
        return st;
    for (size_t k = 0; k < nparts; k++)
        if (!prog_wide_eligible(parts[k].prog)) {
            cft_set_error("%s: part %zu's program is not pure (it uses deposits, activity, "
                          "indexed scratch or flag control); run it with "
                          "cftx_program_run_batch", kWho, k);
            return CFT_ERR_UNSUPPORTED;
        }

    // host banks and host states go into one staging block each, 256-byte aligned
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<const void *> dbank(nparts, nullptr);
    std::vector<uint8_t *> dstate(nparts, nullptr);
    std::vector<cft_buffer *> sbuf(nparts, nullptr);
    size_t bank_off = 0, state_off = 0;
    std::vector<size_t> boff(nparts, SIZE_MAX), soff(nparts, SIZE_MAX);
    for (size_t k = 0; k < nparts; k++) {
        const cftx_batch_part &P = parts[k];
        if (!(P.prog->flags & CFT_PROG_FLAG_BANK_EXT)) {
            dbank[k] = P.prog->d_consts;
        } else if (P.bank_bytes) {
            if (find_buffer(dev, P.bank, P.bank_bytes)) {
                if ((st = bind_in(dev, ROLE_BANK, P.bank, P.bank_bytes, &dbank[k], kWho)) !=
                    CFT_OK)
                    return st;
            } else {
                boff[k] = bank_off;
                bank_off += align(P.bank_bytes);
            }
        }
        if (!P.n)
            continue;
        if (find_buffer(dev, P.state, P.state_bytes)) {
            void *d = nullptr;
            if ((st = bind_out(dev, ROLE_SIN, P.state, P.state_bytes, &d, &sbuf[k], kWho)) !=
                CFT_OK)
                return st;
            dstate[k] = (uint8_t *)d;
        } else {
            soff[k] = state_off;
            state_off += align(P.state_bytes);
        }
    }
    if ((bank_off && (st = arena_reserve(dev, dev->stage[ROLE_BANK], bank_off, kWho)) != CFT_OK) ||
        (state_off &&
         (st = arena_reserve(dev, dev->stage[ROLE_SOUT], state_off, kWho)) != CFT_OK) ||
        (st = arena_reserve(dev, dev->stage[ROLE_LFLAGS], total, kWho)) != CFT_OK)
        return st;
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < nparts && e == hipSuccess; k++) {
        const cftx_batch_part &P = parts[k];
        if (boff[k] != SIZE_MAX) {
            dbank[k] = (uint8_t *)dev->stage[ROLE_BANK].p + boff[k];
            e = hipMemcpyAsync((void *)dbank[k], P.bank, P.bank_bytes, hipMemcpyHostToDevice,
                               dev->stream);
        }
        if (e == hipSuccess && soff[k] != SIZE_MAX) {
            dstate[k] = (uint8_t *)dev->stage[ROLE_SOUT].p + soff[k];
            e = hipMemcpyAsync(dstate[k], P.state, P.state_bytes, hipMemcpyHostToDevice,
                               dev->stream);
        }
    }
    if (e != hipSuccess)
        return dev_fail(dev, e, kWho, "staging banks and states");

    std::vector<MultiPart> mp(nparts);
    for (size_t k = 0; k < nparts; k++) {
        const cftx_batch_part &P = parts[k];
        MultiPart &M = mp[k];
        M.prog = P.prog;
        M.dbank = dbank[k];
        M.width = P.lanes_per_instance;
        M.steps.assign(P.steps, P.steps + P.n);
        M.saddr.resize(P.n);
        const size_t stride = (size_t)P.prog->n_in * P.prog->esz;
        for (size_t i = 0; i < P.n; i++)
            M.saddr[i] = (uint64_t)(uintptr_t)(dstate[k] + i * stride);
    }
    uint8_t *dlf = (uint8_t *)dev->stage[ROLE_LFLAGS].p;
    MultiRun run;
    if ((st = multi_launch(dev, mp, dlf, tune, run, kWho)) != CFT_OK)
        return st;
    std::vector<uint8_t> hlf(run.saddr.size());
    for (size_t k = 0; k < nparts && e == hipSuccess; k++)
        if (soff[k] != SIZE_MAX)
            e = hipMemcpyAsync(parts[k].state, dstate[k], parts[k].state_bytes,
                               hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess && !hlf.empty())
        e = hipMemcpyAsync(hlf.data(), dlf, hlf.size(), hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess && run.launched)
        e = hipMemcpyAsync(dev->hwords, dev->dwords, 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, kWho, "running the launch");
    if (!run.launched)
        dev->hwords[0] = dev->hwords[1] = 0;
    timer_collect(dev);
    for (size_t k = 0; k < nparts; k++) {
        if (sbuf[k])
            sbuf[k]->host_valid = false;
        parts[k].width_used = mp[k].width;
        if (parts[k].lane_flags)
            memset(parts[k].lane_flags, 0, parts[k].n);
    }
    uint32_t bus = 0;
    for (size_t q = 0; q < hlf.size(); q++) {
        const cftx_batch_part &P = parts[run.part_of[q]];
        if (P.lane_flags)
            P.lane_flags[run.lane_of[q]] = hlf[q];
        if (hlf[q] & 0x20u) bus |= CFT_STATUS_DEPOSIT_OVERFLOW;
        if (hlf[q] & 0x40u) bus |= CFT_STATUS_SCRATCH_RANGE;
        if (hlf[q] & 0x80u) bus |= CFT_STATUS_MARKED;
    }
    dev->batch_launches = run.launched ? 1 : 0;
    dev->batch_width = 0;
    cft_emit_flags(dev, dev->hwords[0], flags_out);
    if (bus_out)
        *bus_out = dev->hwords[1] | bus;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_program_prepare(cft_program *prog, uint32_t widths,
                                                   uint32_t *ready)
{
    static const char *who = "cftx_program_prepare";
    cft_clear_error();
    if (ready)
        *ready = 0;
    if (!prog) {
        cft_set_error("%s: prog is NULL", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (widths & ~0x7Fu) {
        cft_set_error("%s: widths 0x%lx names a width past 64 (bit k is width 2^k, k <= 6)", who,
                      (unsigned long)widths);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_status st = dev_enter(prog->dev, who);
    if (st != CFT_OK)
        return st;
    if (!prog_wide_eligible(prog)) {
        cft_set_error("%s: the program is not pure (deposits, activity, indexed scratch or "
                      "flag control); it runs one lane an instance", who);
        return CFT_ERR_UNSUPPORTED;
    }
    // widest first; a width whose wave cannot hold its instances' registers
    // in LDS even at the fewest registers seen is not built (as the planner
    // skips it)
    uint32_t ok = 0, fewest = 0xFFFFFFFFu;
    const uint32_t nq = rows_per_el(prog->fmt);
    for (int k = 6; k >= 0; k--) {
        if (!((widths >> k) & 1u))
            continue;
        const uint32_t ipw = 64u >> k;
        if (fewest != 0xFFFFFFFFu &&
            wave_lds(prog->fmt, (prog->n_consts + fewest * ipw) * nq) > prog->dev->lds_max)
            continue;
        for (uint32_t b : budgets(prog, 1 << k))
            if (WideDev *wd = wide_at(prog, 1 << k, b)) {
                ok |= 1u << k;
                fewest = std::min(fewest, wd->wp.nvregs);
            }
    }
    if (ready)
        *ready = ok;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_multi_last_plan(cft_device *dev, uint64_t out[6])
{
    if (!dev || !out)
        return CFT_ERR_INVALID_ARGUMENT;
    memcpy(out, dev->multi_plan, sizeof dev->multi_plan);
    return CFT_OK;
}
