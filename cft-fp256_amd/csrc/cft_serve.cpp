// The serving runtime (cftx_server_*, include/cft_mi355x.h).
//
// Clients submit jobs from any thread.  One scheduler thread owns the
// workers - each its own handle on the GPU (stream, staging arenas, flag
// words) - and loops:
//  - it completes what has passed its event: an elementwise batch, a whole
//    program batch, or a program slice (each lane's slice trips come off its
//    steps and its flag byte is ORed into its job; a job whose lanes are all
//    done completes);
//  - it gives each free worker its next launch.  An elementwise worker takes
//    every waiting elementwise job.  A program worker admits waiting jobs of
//    pure programs into its running set (oldest first, while the device
//    memory budget allows; a program's jobs gather on the worker that holds
//    its others, a new program goes to the least loaded worker) and launches
//    the next slice over the whole set:
//    one multi-program launch a format, every lane advancing by at most its
//    program's slice trips, a whole number of segments sized so the slice
//    takes about the slice time (measured slices calibrate the estimate).
//    A worker keeps two slices queued on its stream: the next is formed and
//    launched (its lanes' trips committed) while the one before runs, so the
//    GPU does not wait on the host between slices.
//    Jobs of impure programs run whole, one batched launch a program, on
//    the first program worker once it has drained its set.
//
// A job's bits and flags are what it gives alone: elementwise elements are
// independent; a program lane's results depend on its own state and trips,
// and slices cut a lane's steps at whole segments, so the chain of segments
// it runs is the chain it runs alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include <string.h>

#include "seq_prog.hpp"

namespace {

constexpr size_t kMaxElemBatch = 256;
constexpr int kCanElem = 1, kCanProg = 2;
const char *const kWho = "cftx_server";

struct SJob {
    uint64_t ticket = 0;
    cftx_job job;
    std::vector<uint8_t> bank;
    std::vector<uint32_t> steps;
    const void *da = nullptr, *db = nullptr, *dc = nullptr;
    void *dd = nullptr, *dstate = nullptr;
    size_t footprint = 0;
    cftx_job_result res;
    bool done = false;
    bool claimed = false;                // a cftx_server_wait holds this ticket
    // a program job in a running set
    bool pure = false;
    std::vector<uint32_t> left;          // trips to go, a lane
    size_t lanes_left = 0;
    uint32_t fl = 0, bus = 0;
    const void *dbank = nullptr;
};

// An elementwise batch or a whole program batch.
struct Batch {
    std::vector<SJob *> jobs;
    bool program = false;
    std::vector<uint32_t> steps, order;
    std::vector<size_t> lane0;           // first lane of each job
    void *dstate = nullptr;              // gathered states (worker arena)
    uint8_t *dlf = nullptr;              // lane flags
};

// A lane in a slice: its job, its index, whether this slice runs its last
// trips.
struct LaneRef {
    SJob *job;
    uint32_t lane;
    bool last;
};

// One multi-program launch of a slice (one format).
struct SliceRun {
    std::vector<MultiPart> parts;
    std::vector<std::vector<LaneRef>> lanes;   // a part's lanes
    MultiRun run;
    size_t lf0 = 0;                      // its positions' first byte in the lane-flag block
    int fmt = 0;
};

// A slice in flight and the host buffers it holds until it has passed.
struct Slot {
    std::vector<SliceRun> runs;
    hipEvent_t ev = nullptr, t0 = nullptr, t1 = nullptr;
    uint8_t *hlf = nullptr;              // pinned lane flags
    size_t hlf_cap = 0;
    PinnedStage pin[4];                  // a format's tables
};

constexpr int kSlots = 2;                // slices a program worker keeps queued

enum { IDLE = 0, ELEM_BATCH, WHOLE_BATCH };

struct Worker {
    cft_device *dev = nullptr;
    int can = 0;
    int index = 0;
    bool first_prog = false;             // runs the whole batches of impure programs
    hipEvent_t ev = nullptr;             // end of the batch in flight
    int busy = IDLE;                     // a batch in flight (slices are counted apart)
    Batch batch;
    Slot slots[kSlots];
    std::deque<int> inflight;            // slots in flight, oldest first
    std::vector<SJob *> set;             // admitted program jobs
    std::unordered_map<const cft_program *, size_t> held;   // jobs in the set, a program
    bool draining = false;               // admits nothing until its set is empty
    // an elementwise batch's flags: a slot area a job (kernels.h), zero
    // between batches, gathered into the pinned per-job words
    uint32_t *dflags = nullptr;
    uint32_t *hflags = nullptr;          // pinned
    uint32_t *hflags_dev = nullptr;      // hflags' device address
    bool flags_dirty = false;            // a failed batch may have left bits
    uint8_t *hlf = nullptr;              // pinned lane flags of a whole batch
    size_t hlf_cap = 0;
    std::vector<uint8_t *> retired_host; // outgrown pinned buffers, freed at close
    double us_per_bundle[4] = {0, 0, 0, 0};   // measured, a format
    std::unordered_map<const cft_program *, double> body;   // bundles a trip, last slice
};

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Until a slice of the format has been timed: microseconds a bundle.
const double kBundleGuess[4] = {0.15, 0.3, 0.8, 1.4};

} // namespace

struct cftx_server {
    cft_device *owner = nullptr;
    std::vector<Worker> workers;
    std::mutex mu;
    std::condition_variable cv_new, cv_done;
    std::deque<SJob *> incoming;
    std::deque<SJob *> waiting;          // scheduler-owned
    std::unordered_map<uint64_t, SJob *> jobs;
    std::thread sched;
    bool stop = false;
    size_t budget = 0, in_use = 0;
    uint64_t next_ticket = 1, outstanding = 0;
    double t0 = 0;
    cftx_server_stats stats;
    uint32_t max_jobs = 0;               // 0: the defaults
    double slice_ms = 5.0;
    std::map<std::vector<uint8_t>, void *> banks;   // device copies of submitted banks
    // a program's jobs gather on one program worker while it holds any
    std::unordered_map<const cft_program *, int> home;
    // slices leave room on every CU for the elementwise workers' kernels
    MultiTune slice_tune;
};

namespace {

// Under S->mu.
void complete(cftx_server *S, SJob *j, double t)
{
    j->res.done_s = t - S->t0;
    j->done = true;
    S->in_use -= j->footprint;
    S->outstanding--;
}

void lane_byte(SJob *j, uint8_t b)
{
    j->fl |= b & 0x1Fu;
    if (b & 0x20u) j->bus |= CFT_STATUS_DEPOSIT_OVERFLOW;
    if (b & 0x40u) j->bus |= CFT_STATUS_SCRATCH_RANGE;
    if (b & 0x80u) j->bus |= CFT_STATUS_MARKED;
}

void note_launch(cftx_server *S, size_t jobs, bool program)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.batches++;
    if (program) S->stats.program_batches++; else S->stats.elem_batches++;
    S->stats.max_batch_jobs = std::max<uint64_t>(S->stats.max_batch_jobs, jobs);
}

void start_jobs(cftx_server *S, const std::vector<SJob *> &js, size_t batch_jobs)
{
    const double t = now_s() - S->t0;
    for (SJob *j : js) {
        if (j->res.start_s == 0)
            j->res.start_s = t;
        j->res.batch_jobs = std::max<uint32_t>(j->res.batch_jobs, (uint32_t)batch_jobs);
    }
}

// Jobs taken off the waiting list whose memory is held: fail them all.
void fail_jobs(cftx_server *S, std::vector<SJob *> &js, cft_status st)
{
    std::lock_guard<std::mutex> lk(S->mu);
    const double t = now_s();
    for (SJob *j : js) {
        j->res.status = st;
        complete(S, j, t);
    }
    js.clear();
    S->cv_done.notify_all();
}

// ---- elementwise batches and whole program batches ------------------------

void finish_batch(cftx_server *S, Worker &w)
{
    Batch &B = w.batch;
    const double t = now_s();
    std::lock_guard<std::mutex> lk(S->mu);
    for (size_t k = 0; k < B.jobs.size(); k++) {
        SJob *j = B.jobs[k];
        if (!B.program) {
            j->res.flags = w.hflags[k];
        } else {
            for (size_t i = 0; i < j->job.n; i++)
                lane_byte(j, w.hlf[B.lane0[k] + i]);
            j->res.flags = j->fl;
            j->res.bus = j->bus;
        }
        complete(S, j, t);
    }
    w.busy = IDLE;
    B = Batch();
    S->cv_done.notify_all();
}

hipError_t launch_elem(Worker &w)
{
    Batch &B = w.batch;
    cft_device *ex = w.dev;
    constexpr size_t area = (size_t)CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS;
    hipError_t e = plant_launch(ex);
    if (e == hipSuccess && w.flags_dirty) {
        e = cftk_fill32(w.dflags, 0u, kMaxElemBatch * area, ex->stream);
        if (e == hipSuccess)
            w.flags_dirty = false;
    }
    for (size_t k = 0; e == hipSuccess && k < B.jobs.size(); k++) {
        const cftx_job &J = B.jobs[k]->job;
        SJob *j = B.jobs[k];
        const int grid = grid_for(ex, J.n);
        const uint32_t sm = grid > 1 ? CFTK_FLAG_SLOTS - 1 : 0u;
        if (J.kind == CFTX_JOB_ELEM)
            e = cftk_elem(J.fmt, J.op, J.rnd, j->da, j->db, j->dc, j->dd, J.n, 0u,
                          w.dflags + k * area, sm, grid, ex->stream);
        else
            e = cftk_divsqrt(J.fmt, J.kind == CFTX_JOB_DIV ? 0 : 1, J.rnd, j->da, j->db, j->dd,
                             J.n, 0u, w.dflags + k * area, sm, grid, ex->stream);
    }
    if (e == hipSuccess && !B.jobs.empty())
        e = cftk_flags_gather(w.dflags, CFTK_FLAG_SLOTS, w.hflags_dev, (int)B.jobs.size(),
                              ex->stream);
    return e;
}

// Pinned host bytes, grown at least twofold; an outgrown buffer is kept in
// `retired` (freeing pinned memory waits for the device) until close.
cft_status pinned_bytes(uint8_t *&p, size_t &cap, size_t n, std::vector<uint8_t *> &retired)
{
    if (cap >= n)
        return CFT_OK;
    const size_t want = std::max<size_t>({n, cap * 2, (size_t)1 << 16});
    uint8_t *q = nullptr;
    if (hipHostMalloc((void **)&q, want) != hipSuccess) {
        (void)hipGetLastError();
        cft_set_error("%s: %zu bytes of pinned host memory could not be allocated", kWho, want);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    if (p)
        retired.push_back(p);
    p = q;
    cap = want;
    return CFT_OK;
}

// Jobs of one impure program, run whole: states gathered into the worker's
// arena, one persistent batched launch, states scattered back.
cft_status launch_whole(Worker &w)
{
    Batch &B = w.batch;
    cft_device *ex = w.dev;
    cft_program *prog = B.jobs[0]->job.prog;
    const size_t esz = prog->esz, slots = prog->n_in;
    size_t n = 0;
    for (SJob *j : B.jobs) {
        B.lane0.push_back(n);
        n += j->job.n;
    }
    for (SJob *j : B.jobs)
        B.steps.insert(B.steps.end(), j->steps.begin(), j->steps.end());
    B.order.resize(n);
    for (size_t i = 0; i < n; i++)
        B.order[i] = (uint32_t)i;
    std::stable_sort(B.order.begin(), B.order.end(),
                     [&](uint32_t x, uint32_t y) { return B.steps[x] > B.steps[y]; });
    cft_status st = arena_grow(ex, ex->stage[ROLE_SIN], n * slots * esz, kWho);
    if (st == CFT_OK)
        st = arena_grow(ex, ex->stage[ROLE_LFLAGS], n, kWho);
    if (st == CFT_OK)
        st = pinned_bytes(w.hlf, w.hlf_cap, n, w.retired_host);
    if (st != CFT_OK)
        return st;
    B.dstate = ex->stage[ROLE_SIN].p;
    B.dlf = (uint8_t *)ex->stage[ROLE_LFLAGS].p;
    hipError_t e = hipSuccess;
    for (size_t k = 0; e == hipSuccess && k < B.jobs.size(); k++)
        e = hipMemcpyAsync((uint8_t *)B.dstate + B.lane0[k] * slots * esz, B.jobs[k]->dstate,
                           B.jobs[k]->job.n * slots * esz, hipMemcpyDeviceToDevice, ex->stream);
    if (e != hipSuccess)
        return dev_fail(ex, e, kWho, "gathering states");
    BatchIO io;
    io.bank = B.jobs[0]->dbank;
    io.state = B.dstate;
    io.lflags = B.dlf;
    io.steps = B.steps.data();
    io.order = B.order.data();
    io.n = n;
    int width = 0;
    st = batch_persistent(ex, prog, io, kWho, &width);
    if (st != CFT_OK)
        return st;
    for (size_t k = 0; e == hipSuccess && k < B.jobs.size(); k++)
        e = hipMemcpyAsync(B.jobs[k]->dstate, (uint8_t *)B.dstate + B.lane0[k] * slots * esz,
                           B.jobs[k]->job.n * slots * esz, hipMemcpyDeviceToDevice, ex->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w.hlf, B.dlf, n, hipMemcpyDeviceToHost, ex->stream);
    if (e != hipSuccess)
        return dev_fail(ex, e, kWho, "scattering states");
    return CFT_OK;
}

// ---- program slices ---------------------------------------------------------

// The device copy of a job's constants: the image's own, or the submitted
// bank uploaded once per distinct content.
cft_status job_bank(cftx_server *S, SJob *j)
{
    cft_program *p = j->job.prog;
    if (!(p->flags & CFT_PROG_FLAG_BANK_EXT)) {
        j->dbank = p->d_consts;
        return CFT_OK;
    }
    if (j->bank.empty()) {
        j->dbank = nullptr;
        return CFT_OK;
    }
    auto it = S->banks.find(j->bank);
    if (it == S->banks.end()) {
        void *d = nullptr;
        if (hipMalloc(&d, j->bank.size()) != hipSuccess ||
            hipMemcpy(d, j->bank.data(), j->bank.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            if (d)
                (void)hipFree(d);
            cft_set_error("%s: a constant bank could not be copied to the device", kWho);
            return CFT_ERR_OUT_OF_MEMORY;
        }
        it = S->banks.emplace(j->bank, d).first;
    }
    j->dbank = it->second;
    return CFT_OK;
}

void release(cftx_server *S, Worker &w, const cft_program *p)
{
    auto it = w.held.find(p);
    if (it != w.held.end() && --it->second == 0) {
        w.held.erase(it);
        auto h = S->home.find(p);
        if (h != S->home.end() && h->second == w.index)
            S->home.erase(h);
    }
}

// Bundles of work a program worker's set still holds (an estimate).
double set_load(const Worker &w)
{
    double load = 0;
    for (const SJob *j : w.set) {
        auto b = w.body.find(j->job.prog);
        load += (double)j->lanes_left * (b != w.body.end() ? b->second : 20.0);
    }
    return load;
}

// The oldest slice in flight has passed: flags into jobs, lanes that ran
// their last trips off their jobs, finished jobs completed.
void finish_slot(cftx_server *S, Worker &w)
{
    Slot &sl = w.slots[w.inflight.front()];
    w.inflight.pop_front();
    float ms = 0.f;
    const bool timed = hipEventElapsedTime(&ms, sl.t0, sl.t1) == hipSuccess;
    if (!timed)
        (void)hipGetLastError();
    double span = 0;
    int fmt = 0;
    for (SliceRun &R : sl.runs) {
        span += R.run.makespan;
        fmt = R.fmt;
        for (size_t q = 0; q < R.run.saddr.size(); q++) {
            const LaneRef &L = R.lanes[R.run.part_of[q]][R.run.lane_of[q]];
            lane_byte(L.job, sl.hlf[R.lf0 + q]);
            if (L.last)
                L.job->lanes_left--;
        }
        for (size_t p = 0; p < R.parts.size(); p++)
            if (R.run.body[p] > 0)
                w.body[R.parts[p].prog] = R.run.body[p];
    }
    // a slice of one format (most are) calibrates that format's estimate
    if (timed && span > 0 && sl.runs.size() == 1) {
        const double v = ms * 1e3 / span;
        double &u = w.us_per_bundle[fmt];
        u = u > 0 ? 0.7 * u + 0.3 * v : v;
    }
    sl.runs.clear();
    const double t = now_s();
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<SJob *> keep;
    for (SJob *j : w.set) {
        if (j->lanes_left == 0) {
            j->res.flags = j->fl;
            j->res.bus = j->bus;
            release(S, w, j->job.prog);
            complete(S, j, t);
        } else {
            keep.push_back(j);
        }
    }
    w.set.swap(keep);
    S->cv_done.notify_all();
}

// Trips a lane of `p` advances in this slice.
uint32_t slice_trips(cftx_server *S, Worker &w, cft_program *p)
{
    if (S->max_jobs == 1 || S->slice_ms <= 0)
        return 0xFFFFFFFFu;
    const uint32_t seg = std::max<uint32_t>(1, p->code[p->top_repeat].imm);
    auto it = w.body.find(p);
    if (it == w.body.end())
        return seg;                      // one segment, to learn the program's cost
    const double us = w.us_per_bundle[p->fmt] > 0 ? w.us_per_bundle[p->fmt]
                                                   : kBundleGuess[p->fmt];
    const double k = S->slice_ms * 1e3 / ((double)seg * it->second * us);
    const double trips = std::max(1.0, std::floor(k)) * seg;
    return (uint32_t)std::min(trips, 2147483648.0);
}

// The next slice of w's set into a free slot: each lane with trips not yet
// launched takes up to its program's slice trips.  *launched is false when
// every lane's trips are already in flight.
cft_status launch_slice(cftx_server *S, Worker &w, bool *launched)
{
    *launched = false;
    cft_device *ex = w.dev;
    int k = 0;
    while (k < kSlots && std::find(w.inflight.begin(), w.inflight.end(), k) != w.inflight.end())
        k++;
    Slot &sl = w.slots[k];
    sl.runs.clear();
    std::map<int, size_t> by_fmt;
    std::map<std::pair<const cft_program *, const void *>, size_t> part_of;
    std::unordered_map<const cft_program *, uint32_t> trips;
    size_t lanes = 0, jobs = 0;
    for (SJob *j : w.set) {
        cft_program *p = j->job.prog;
        bool any = false;
        for (uint32_t i = 0; i < (uint32_t)j->left.size() && !any; i++)
            any = j->left[i] != 0;
        if (!any)
            continue;
        jobs++;
        auto f = by_fmt.find(p->fmt);
        if (f == by_fmt.end()) {
            f = by_fmt.emplace(p->fmt, sl.runs.size()).first;
            sl.runs.emplace_back();
            sl.runs.back().fmt = p->fmt;
        }
        SliceRun &R = sl.runs[f->second];
        const auto key = std::make_pair((const cft_program *)p, j->dbank);
        auto pk = part_of.find(key);
        if (pk == part_of.end()) {
            pk = part_of.emplace(key, R.parts.size()).first;
            R.parts.emplace_back();
            R.parts.back().prog = p;
            R.parts.back().dbank = j->dbank;
            R.lanes.emplace_back();
        }
        MultiPart &M = R.parts[pk->second];
        auto &L = R.lanes[pk->second];
        auto tr = trips.find(p);
        if (tr == trips.end())
            tr = trips.emplace(p, slice_trips(S, w, p)).first;
        const size_t stride = (size_t)p->n_in * p->esz;
        for (uint32_t i = 0; i < (uint32_t)j->left.size(); i++) {
            if (!j->left[i])
                continue;
            const uint32_t t = std::min(j->left[i], tr->second);
            j->left[i] -= t;
            M.saddr.push_back((uint64_t)(uintptr_t)((uint8_t *)j->dstate + i * stride));
            M.steps.push_back(t);
            L.push_back(LaneRef{j, i, j->left[i] == 0});
            lanes++;
        }
    }
    if (lanes == 0)
        return CFT_OK;
    cft_status st = arena_grow(ex, ex->stage[ROLE_LFLAGS], lanes, kWho);
    if (st == CFT_OK)
        st = pinned_bytes(sl.hlf, sl.hlf_cap, lanes, w.retired_host);
    if (st != CFT_OK)
        return st;
    uint8_t *dlf = (uint8_t *)ex->stage[ROLE_LFLAGS].p;
    hipError_t e = hipEventRecord(sl.t0, ex->stream);
    size_t off = 0;
    for (SliceRun &R : sl.runs) {
        if (e != hipSuccess)
            break;
        R.lf0 = off;
        if ((st = multi_launch(ex, R.parts, dlf + off, S->slice_tune, R.run, kWho,
                               &sl.pin[R.fmt])) != CFT_OK)
            return st;
        off += R.run.saddr.size();
    }
    if (e == hipSuccess)
        e = hipEventRecord(sl.t1, ex->stream);
    if (e == hipSuccess && off)
        e = hipMemcpyAsync(sl.hlf, dlf, off, hipMemcpyDeviceToHost, ex->stream);
    if (e == hipSuccess)
        e = hipEventRecord(sl.ev, ex->stream);
    if (e != hipSuccess)
        return dev_fail(ex, e, kWho, "launching a slice");
    w.inflight.push_back(k);
    *launched = true;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stats.slices++;
        S->stats.max_slice_lanes = std::max<uint64_t>(S->stats.max_slice_lanes, lanes);
    }
    note_launch(S, jobs, true);
    start_jobs(S, w.set, jobs);
    return CFT_OK;
}

// Admit waiting jobs of pure programs into w's set, oldest first, within the
// memory budget (in order: a job that does not fit holds the ones behind it).
void admit(cftx_server *S, Worker &w)
{
    const bool plain = S->max_jobs == 1;
    if (plain && !w.set.empty())
        return;
    for (auto it = S->waiting.begin(); it != S->waiting.end();) {
        SJob *j = *it;
        if (j->job.kind != CFTX_JOB_PROGRAM || !j->pure) {
            ++it;
            continue;
        }
        if (S->in_use + j->footprint > S->budget && S->in_use > 0)
            break;
        if (S->max_jobs > 1 && w.set.size() >= S->max_jobs)
            break;
        // the program's home worker takes it; a program with none goes to the
        // least loaded program worker
        const cft_program *p = j->job.prog;
        auto h = S->home.find(p);
        int target = w.index;
        if (h != S->home.end()) {
            target = h->second;
        } else if (!plain) {
            double least = set_load(w);
            for (const Worker &o : S->workers)
                if ((o.can & kCanProg) && o.index != w.index && !o.draining &&
                    set_load(o) < least * 0.999) {
                    least = set_load(o);
                    target = o.index;
                }
        }
        if (target != w.index) {
            ++it;
            continue;
        }
        it = S->waiting.erase(it);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->in_use += j->footprint;
            S->stats.peak_bytes = std::max<uint64_t>(S->stats.peak_bytes, S->in_use);
        }
        const cft_status st = job_bank(S, j);
        if (st != CFT_OK) {
            std::vector<SJob *> one{j};
            fail_jobs(S, one, st);
            continue;
        }
        j->left = j->steps;
        j->lanes_left = 0;
        for (uint32_t s : j->left)
            j->lanes_left += s ? 1 : 0;
        if (j->lanes_left == 0) {                 // nothing to run: done as it stands
            std::lock_guard<std::mutex> lk(S->mu);
            j->res.start_s = now_s() - S->t0;
            complete(S, j, now_s());
            S->cv_done.notify_all();
        } else {
            w.set.push_back(j);
            w.held[p]++;
            S->home[p] = w.index;
        }
        if (plain)
            break;
    }
}

// A failed slice fails every job of the set (the stream drained first).
void drop_set(cftx_server *S, Worker &w, cft_status st)
{
    (void)hipStreamSynchronize(w.dev->stream);
    (void)hipGetLastError();
    for (int k : w.inflight)
        w.slots[k].runs.clear();
    w.inflight.clear();
    for (SJob *j : w.set)
        release(S, w, j->job.prog);
    fail_jobs(S, w.set, st);
}

// ---- dispatch ---------------------------------------------------------------

bool elem_waiting(const cftx_server *S)
{
    for (const SJob *j : S->waiting)
        if (j->job.kind != CFTX_JOB_PROGRAM)
            return true;
    return false;
}

void start_elem(cftx_server *S, Worker &w)
{
    Batch B;
    const size_t cap = S->max_jobs ? S->max_jobs : kMaxElemBatch;
    for (auto it = S->waiting.begin(); it != S->waiting.end() && B.jobs.size() < cap;) {
        if ((*it)->job.kind != CFTX_JOB_PROGRAM) {
            B.jobs.push_back(*it);
            it = S->waiting.erase(it);
        } else {
            ++it;
        }
    }
    if (B.jobs.empty())
        return;
    w.batch = std::move(B);
    note_launch(S, w.batch.jobs.size(), false);
    start_jobs(S, w.batch.jobs, w.batch.jobs.size());
    hipError_t e = launch_elem(w);
    if (e == hipSuccess)
        e = hipEventRecord(w.ev, w.dev->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail_jobs(S, w.batch.jobs, CFT_ERR_INTERNAL);
        w.batch = Batch();
        w.flags_dirty = true;                    // kernels may have run without the gather
        return;
    }
    w.busy = ELEM_BATCH;
}

// A whole batch: waiting jobs of the oldest waiting impure program.
void start_whole(cftx_server *S, Worker &w, SJob *first)
{
    Batch B;
    B.program = true;
    size_t lanes = 0;
    const cft_program *prog = first->job.prog;
    const std::vector<uint8_t> bank = first->bank;
    for (auto it = S->waiting.begin(); it != S->waiting.end();) {
        SJob *j = *it;
        if (S->max_jobs && B.jobs.size() >= S->max_jobs)
            break;
        // the first job is always taken, whatever its size; later ones join
        // while the batch stays within 2^22 lanes and the memory budget
        if (j->job.kind == CFTX_JOB_PROGRAM && j->job.prog == prog && j->bank == bank &&
            (B.jobs.empty() ||
             (lanes + j->job.n <= (1u << 22) && S->in_use + j->footprint <= S->budget))) {
            B.jobs.push_back(j);
            lanes += j->job.n;
            it = S->waiting.erase(it);
            std::lock_guard<std::mutex> lk(S->mu);
            S->in_use += j->footprint;
            S->stats.peak_bytes = std::max<uint64_t>(S->stats.peak_bytes, S->in_use);
        } else {
            ++it;
        }
    }
    if (B.jobs.empty())
        return;
    w.batch = std::move(B);
    cft_status st = job_bank(S, w.batch.jobs[0]);
    for (SJob *j : w.batch.jobs)
        j->dbank = w.batch.jobs[0]->dbank;
    note_launch(S, w.batch.jobs.size(), true);
    start_jobs(S, w.batch.jobs, w.batch.jobs.size());
    hipError_t e = hipSuccess;
    if (st == CFT_OK)
        st = launch_whole(w);
    if (st == CFT_OK)
        e = hipEventRecord(w.ev, w.dev->stream);
    if (st != CFT_OK || e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(w.dev->stream);
        fail_jobs(S, w.batch.jobs, st != CFT_OK ? st : CFT_ERR_INTERNAL);
        w.batch = Batch();
        return;
    }
    w.busy = WHOLE_BATCH;
}

// A program worker with room for a slice: a whole batch for a waiting
// impure program once drained, else admissions and the next slice.
void start_program(cftx_server *S, Worker &w)
{
    SJob *impure = nullptr;
    for (SJob *j : S->waiting)
        if (j->job.kind == CFTX_JOB_PROGRAM && !j->pure) {
            impure = j;
            break;
        }
    if (impure && w.first_prog)
        w.draining = true;
    if (w.draining && w.set.empty() && w.inflight.empty()) {
        w.draining = false;
        if (impure) {
            if (S->in_use + impure->footprint > S->budget && S->in_use > 0) {
                w.draining = true;       // wait for memory
                return;
            }
            start_whole(S, w, impure);
            return;
        }
    }
    if (!w.draining)
        admit(S, w);
    if (w.set.empty())
        return;
    bool launched = false;
    const double t0 = now_s();
    const cft_status st = launch_slice(S, w, &launched);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stats.slice_host_s += now_s() - t0;
    }
    if (st != CFT_OK)
        drop_set(S, w, st);
}

// Can w take a launch now?
bool has_room(const Worker &w)
{
    if (w.busy != IDLE)
        return false;
    return (int)w.inflight.size() < kSlots;
}

void dispatch(cftx_server *S, Worker &w)
{
    if (w.can == kCanElem) {
        start_elem(S, w);
        return;
    }
    if (w.can == kCanProg) {
        start_program(S, w);
        return;
    }
    // one worker for both: the oldest waiting job decides; an elementwise
    // batch waits for the slices in flight to pass
    if (!S->waiting.empty() && S->waiting.front()->job.kind != CFTX_JOB_PROGRAM) {
        if (w.inflight.empty())
            start_elem(S, w);
        return;
    }
    start_program(S, w);
    if (w.busy == IDLE && w.inflight.empty() && elem_waiting(S))
        start_elem(S, w);
}

void scheduler(cftx_server *S)
{
    (void)hipSetDevice(S->owner->hip_index);
    for (;;) {
        bool idle_pass = true;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            while (!S->incoming.empty()) {
                S->waiting.push_back(S->incoming.front());
                S->incoming.pop_front();
            }
            if (S->stop && S->waiting.empty()) {
                bool any = false;
                for (Worker &w : S->workers)
                    any |= w.busy != IDLE || !w.set.empty() || !w.inflight.empty();
                if (!any)
                    return;
            }
        }
        for (Worker &w : S->workers) {
            while (!w.inflight.empty()) {
                const hipError_t q = hipEventQuery(w.slots[w.inflight.front()].ev);
                if (q == hipSuccess) {
                    finish_slot(S, w);
                    idle_pass = false;
                    continue;
                }
                if (q != hipErrorNotReady) {
                    (void)hipGetLastError();
                    drop_set(S, w, CFT_ERR_INTERNAL);
                    idle_pass = false;
                }
                break;
            }
            if (w.busy == IDLE)
                continue;
            const hipError_t q = hipEventQuery(w.ev);
            if (q == hipSuccess) {
                finish_batch(S, w);
                idle_pass = false;
            } else if (q != hipErrorNotReady) {
                (void)hipGetLastError();
                fail_jobs(S, w.batch.jobs, CFT_ERR_INTERNAL);
                w.batch = Batch();
                w.busy = IDLE;
                idle_pass = false;
            }
        }
        for (Worker &w : S->workers) {
            if (!has_room(w))
                continue;
            const size_t before = S->waiting.size() + w.set.size() + w.inflight.size();
            dispatch(S, w);
            if (w.busy != IDLE ||
                S->waiting.size() + w.set.size() + w.inflight.size() != before)
                idle_pass = false;
        }
        if (idle_pass) {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_new.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

// The device memory a program job holds while it runs (its lanes' share of
// the scratch pool and of a gathered state block), an estimate for admission.
size_t program_footprint(const SJob *j)
{
    const cft_program *p = j->job.prog;
    const size_t per = (size_t)seq_scratch_reach(p) * p->esz + (size_t)p->n_in * p->esz + 16;
    return j->job.n * per * 2;
}

} // namespace

extern "C" CFT_API cft_status cftx_server_open(cft_device *dev, int workers, size_t hbm_budget,
                                               cftx_server **out)
{
    cft_clear_error();
    if (!dev || !out) {
        cft_set_error("cftx_server_open: %s is NULL", !dev ? "dev" : "out");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (workers == 0)
        workers = 4;
    if (workers < 1 || workers > 8) {
        cft_set_error("cftx_server_open: %d workers; 1 to 8", workers);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_status st = dev_enter(dev, "cftx_server_open");
    if (st != CFT_OK)
        return st;
    cftx_server *S = new (std::nothrow) cftx_server();
    if (!S)
        return CFT_ERR_OUT_OF_MEMORY;
    S->owner = dev;
    S->budget = hbm_budget ? hbm_budget : ((size_t)8 << 30);
    S->t0 = now_s();
    memset(&S->stats, 0, sizeof S->stats);
    S->workers.resize(workers);
    const int n_elem = workers == 1 ? 0 : workers / 2;
    for (int k = 0; k < workers; k++) {
        Worker &w = S->workers[k];
        w.index = k;
        w.can = workers == 1 ? (kCanElem | kCanProg) : (k < n_elem ? kCanElem : kCanProg);
        w.first_prog = workers == 1 || k == n_elem;
        // every worker on a hardware queue of its own: an elementwise launch
        // never queues behind a program slice, and two program workers'
        // slices run side by side whatever other streams the process holds
        st = device_open(dev->hip_index, dev->scratch_depth, &w.dev, true);
        hipError_t e = hipSuccess;
        if (st == CFT_OK) e = hipEventCreateWithFlags(&w.ev, hipEventDisableTiming);
        for (Slot &sl : w.slots) {
            if (st == CFT_OK && e == hipSuccess)
                e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
            if (st == CFT_OK && e == hipSuccess) e = hipEventCreate(&sl.t0);
            if (st == CFT_OK && e == hipSuccess) e = hipEventCreate(&sl.t1);
        }
        const size_t fbytes = kMaxElemBatch * CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS * 4;
        if (st == CFT_OK && e == hipSuccess) e = hipMalloc(&w.dflags, fbytes);
        if (st == CFT_OK && e == hipSuccess) e = hipMemsetAsync(w.dflags, 0, fbytes, w.dev->stream);
        if (st == CFT_OK && e == hipSuccess) e = hipStreamSynchronize(w.dev->stream);
        if (st == CFT_OK && e == hipSuccess)
            e = hipHostMalloc((void **)&w.hflags, kMaxElemBatch * 4);
        if (st == CFT_OK && e == hipSuccess)
            e = hipHostGetDevicePointer((void **)&w.hflags_dev, w.hflags, 0);
        if (st == CFT_OK && e != hipSuccess) {
            (void)hipGetLastError();
            st = CFT_ERR_OUT_OF_MEMORY;
        }
        if (st != CFT_OK) {
            cftx_server_close(S);
            return st;
        }
    }
    // a CU holds about 12 sequencer waves (registers); the program workers'
    // slices share 8 of them and leave the rest to elementwise launches,
    // which would otherwise wait for a slice to end before they start
    const int n_prog = workers - n_elem;
    if (n_elem > 0)
        S->slice_tune.max_waves_per_cu = (uint32_t)std::max(1, 8 / n_prog);
    S->sched = std::thread(scheduler, S);
    *out = S;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_server_submit(cftx_server *S, const cftx_job *job,
                                                 uint64_t *ticket)
{
    static const char *who = "cftx_server_submit";
    cft_clear_error();
    if (!S || !job || !ticket) {
        cft_set_error("%s: a NULL argument", who);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (job->struct_size != sizeof(cftx_job)) {
        cft_set_error("%s was given a %zu-byte cftx_job and this library's is %zu", who,
                      job->struct_size, sizeof(cftx_job));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_device *dev = S->owner;
    SJob *j = new (std::nothrow) SJob();
    if (!j)
        return CFT_ERR_OUT_OF_MEMORY;
    j->job = *job;
    memset(&j->res, 0, sizeof j->res);
    j->res.struct_size = sizeof j->res;
    // operands: windows of current resident buffers, or of device memory of
    // the server's GPU that the caller allocated (aligned for the kernels'
    // loads, inside one allocation)
    auto window = [&](const void *p, size_t bytes, const char *what) -> void * {
        cft_buffer *b = find_buffer(dev, p, bytes);
        if (!b) {
            const size_t esz = job->kind == CFTX_JOB_PROGRAM && job->prog
                                   ? job->prog->esz
                                   : cft_format_size((cft_format)job->fmt);
            if (device_window(who, dev, what, p, bytes, esz < 16 ? esz : 16) != CFT_OK)
                return nullptr;
            return const_cast<void *>(p);
        }
        if (!b->dev_valid) {
            cft_set_error("%s: %s's buffer is not current on the device; call "
                          "cft_buffer_to_device first", who, what);
            return nullptr;
        }
        return (uint8_t *)b->devp + ((const uint8_t *)p - (const uint8_t *)b->host);
    };
    auto mark_out = [&](void *p, size_t bytes) {
        if (cft_buffer *b = find_buffer(dev, p, bytes))
            b->host_valid = false;
    };
    cft_status st = CFT_OK;
    const cftx_job &J = *job;
    if (J.kind == CFTX_JOB_ELEM || J.kind == CFTX_JOB_DIV || J.kind == CFTX_JOB_SQRT) {
        if (J.fmt < 0 || J.fmt > 3 || J.rnd < 0 || J.rnd > 4 || J.n == 0 || !J.a || !J.d ||
            (J.kind == CFTX_JOB_DIV && !J.b) ||
            (J.kind == CFTX_JOB_ELEM && (J.op < 0 || J.op > 255 || J.op == 24 || J.op == 25 ||
                                         J.op == 28 || J.op == 29 || J.op == 31))) {
            cft_set_error("%s: an elementwise job needs a format 0..3, a rounding 0..4, "
                          "n > 0, a and d (and b for a division), and a non-reduction "
                          "opcode", who);
            st = CFT_ERR_INVALID_ARGUMENT;
        } else {
            const size_t bytes = J.n * cft_format_size((cft_format)J.fmt);
            j->da = window(J.a, bytes, "a");
            j->db = J.b ? window(J.b, bytes, "b") : nullptr;
            j->dc = J.c ? window(J.c, bytes, "c") : nullptr;
            j->dd = window(J.d, bytes, "d");
            if (!j->da || !j->dd || (J.b && !j->db) || (J.c && !j->dc))
                st = CFT_ERR_INVALID_ARGUMENT;
            else
                mark_out(J.d, bytes);
            // a NULL operand the opcode reads is refused exactly as cft_run
            // refuses it (the same table of which operands an opcode reads)
            const unsigned reads = cft_op_reads(J.op);
            if (st == CFT_OK && J.kind == CFTX_JOB_ELEM &&
                (((reads & 2u) && !J.b) || ((reads & 4u) && !J.c))) {
                cft_set_error("%s: opcode %d (%s) reads operand %c, which is NULL", who, J.op,
                              cft_op_name((cft_op)J.op), ((reads & 2u) && !J.b) ? 'b' : 'c');
                st = CFT_ERR_INVALID_ARGUMENT;
            }
        }
    } else if (J.kind == CFTX_JOB_PROGRAM) {
        cft_program *p = J.prog;
        if (!p || p->dev != dev || !(p->flags & CFT_PROG_FLAG_SCRATCH_IO) || p->n_in == 0 ||
            p->n_in != p->n_out || p->max_deposits || p->top_repeat == 0xFFFFFFFFu ||
            J.n == 0 || J.n > 0xFFFFFFFFull || !J.steps || !J.state ||
            J.state_bytes != J.n * p->n_in * p->esz) {
            cft_set_error("%s: a program job needs a segment program loaded on the server's "
                          "device (state as scratch I/O, no deposits, a top-level REPEAT), "
                          "n > 0 lanes, steps and a state of n x slots elements", who);
            st = CFT_ERR_INVALID_ARGUMENT;
        } else {
            st = seq_check_bank(p, J.bank, J.bank_bytes, who);
            if (st == CFT_OK) {
                j->dstate = window(J.state, J.state_bytes, "state");
                if (!j->dstate)
                    st = CFT_ERR_INVALID_ARGUMENT;
                else
                    mark_out(J.state, J.state_bytes);
            }
            if (st == CFT_OK) {
                j->steps.assign(J.steps, J.steps + J.n);
                if (J.bank_bytes)
                    j->bank.assign((const uint8_t *)J.bank,
                                   (const uint8_t *)J.bank + J.bank_bytes);
                j->footprint = program_footprint(j);
                j->pure = prog_wide_eligible(p);
            }
        }
    } else {
        cft_set_error("%s: job kind %d is not CFTX_JOB_ELEM/DIV/SQRT/PROGRAM", who, J.kind);
        st = CFT_ERR_INVALID_ARGUMENT;
    }
    if (st != CFT_OK) {
        delete j;
        return st;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        j->ticket = S->next_ticket++;
        j->res.submit_s = now_s() - S->t0;
        S->jobs[j->ticket] = j;
        S->incoming.push_back(j);
        S->outstanding++;
        S->stats.jobs++;
    }
    S->cv_new.notify_one();
    *ticket = j->ticket;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_server_wait(cftx_server *S, uint64_t ticket,
                                               cftx_job_result *out)
{
    cft_clear_error();
    if (!S || !out) {
        cft_set_error("cftx_server_wait: a NULL argument");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end() || it->second->claimed) {
        cft_set_error(it == S->jobs.end()
                          ? "cftx_server_wait: ticket %llu is not an outstanding job"
                          : "cftx_server_wait: ticket %llu is already being waited for",
                      (unsigned long long)ticket);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    // the job stays in the table while this thread sleeps (submissions may
    // rehash it, so no iterator is kept across the wait); a second waiter
    // on the same ticket is refused above instead of touching a freed job
    SJob *j = it->second;
    j->claimed = true;
    S->cv_done.wait(lk, [&] { return j->done; });
    const size_t want = std::min(out->struct_size ? out->struct_size : sizeof *out,
                                 sizeof(cftx_job_result));
    cftx_job_result r = j->res;
    r.struct_size = want;
    memcpy(out, &r, want);
    S->jobs.erase(ticket);
    delete j;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_server_set_batching(cftx_server *S, uint32_t max_jobs)
{
    cft_clear_error();
    if (!S || max_jobs > kMaxElemBatch) {
        cft_set_error("cftx_server_set_batching: %s", !S ? "srv is NULL" : "at most 256 jobs");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->max_jobs = max_jobs;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_server_set_slice(cftx_server *S, double ms)
{
    cft_clear_error();
    if (!S || !(ms >= 0) || ms > 60000) {
        cft_set_error("cftx_server_set_slice: %s", !S ? "srv is NULL"
                                                      : "the slice time is 0 to 60,000 ms");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->slice_ms = ms;
    return CFT_OK;
}

extern "C" CFT_API double cftx_server_clock(cftx_server *S)
{
    return S ? now_s() - S->t0 : 0.0;
}

extern "C" CFT_API cft_status cftx_server_stats_get(cftx_server *S, cftx_server_stats *out)
{
    if (!S || !out)
        return CFT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    cftx_server_stats s = S->stats;
    const size_t want = std::min(out->struct_size ? out->struct_size : sizeof s, sizeof s);
    s.struct_size = want;
    memcpy(out, &s, want);
    return CFT_OK;
}

extern "C" CFT_API void cftx_server_close(cftx_server *S)
{
    if (!S)
        return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_new.notify_all();
    if (S->sched.joinable())
        S->sched.join();
    for (Worker &w : S->workers) {
        if (w.ev) (void)hipEventDestroy(w.ev);
        for (Slot &sl : w.slots) {
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.t0) (void)hipEventDestroy(sl.t0);
            if (sl.t1) (void)hipEventDestroy(sl.t1);
            if (sl.hlf) (void)hipHostFree(sl.hlf);
        }
        if (w.dflags) (void)hipFree(w.dflags);
        if (w.hflags) (void)hipHostFree(w.hflags);
        if (w.hlf) (void)hipHostFree(w.hlf);
        for (uint8_t *q : w.retired_host)
            (void)hipHostFree(q);
        if (w.dev) cft_close(w.dev);
    }
    for (auto &kv : S->banks)
        (void)hipFree(kv.second);
    for (auto &kv : S->jobs)
        delete kv.second;
    delete S;
}
