// The loaded program (cft_program) and what the program module shares with
// the batched runner (cft_batch.cpp).
#pragma once

#include <vector>

#include <map>

#include "cft_internal.hpp"
#include "kernels.h"
#include "seq_wide.hpp"

// A wide schedule on the device.
struct WideDev {
    WideProgram wp;
    WideRec *d_recs = nullptr;
    bool failed = false;            // the program does not schedule at this width
};

struct cft_program {
    cft_device *dev = nullptr;
    int fmt = 0;
    size_t esz = 0;
    uint32_t max_deposits = 0, n_insns = 0, n_consts = 0, flags = 0;
    uint32_t n_in = 0, n_out = 0, scratch_used = 0;
    bool scratch_active = false;
    uint32_t depth = 0;
    int log2depth = 0;
    std::vector<uint8_t> image;
    std::vector<uint64_t> insns;
    // the decoded form
    std::vector<SeqInsn> code;
    uint32_t nregs = 0;
    int phys_abc[3] = {-1, -1, -1};
    uint32_t top_repeat = 0xFFFFFFFFu;   // the first REPEAT at nesting depth 0
    SeqInsn *d_code = nullptr;
    uint32_t *d_consts = nullptr;         // an image's own constants, on the device
    // the batched runner's captured graph and the launch it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey[24] = {};
    // wide schedules by (width << 32 | scheduling window), built on first use
    std::map<uint64_t, WideDev> wide;
    int wide_eligible = -1;          // -1 not yet known
};

// A persistent batched launch's inputs: device pointers (state and lane
// flags written in place), host steps and lane order (alive until the
// stream passes the launch).
struct BatchIO {
    const void *bank = nullptr, *a = nullptr, *b = nullptr, *c = nullptr;
    void *state = nullptr;
    uint8_t *lflags = nullptr;
    const uint32_t *steps = nullptr;
    const uint32_t *order = nullptr;
    size_t n = 0;
    uint32_t lanes_per_wave = 0, waves_per_group = 0, lanes_per_instance = 0;
    bool graph = false;
};

// The persistent batched runner enqueued on ex's stream (cft_batch.cpp).
cft_status batch_persistent(cft_device *ex, cft_program *prog, const BatchIO &io,
                            const char *who, int *width_out);

// The wide schedule of `prog` at `width` within `max_vregs` registers an
// instance (0: the most half the device's LDS holds for a wave's instances),
// compiled and uploaded on first use (the device current); NULL when the
// program does not schedule there.
WideDev *wide_at(cft_program *prog, int width, uint32_t max_vregs = 0);

// The register budget wide_at uses for max_vregs 0.
uint32_t wide_default_budget(const cft_program *prog, int width);

// Whether the program runs wide (pure: ALU, static scratch, loops).
bool prog_wide_eligible(cft_program *prog);

// One program's share of a multi-program launch.
struct MultiPart {
    cft_program *prog = nullptr;
    const void *dbank = nullptr;      // its constants on the device, or NULL
    std::vector<uint64_t> saddr;      // each lane's state (device address)
    std::vector<uint32_t> steps;      // each lane's trips
    uint32_t width = 0;               // lanes an instance takes: 0 for the planner's
                                      // choice; on return, the width used
};

// A multi-program launch's host side, alive until the stream passes it:
// positions are the lanes with trips, each part's longest first.
struct MultiRun {
    std::vector<SeqMProg> progs;
    std::vector<uint32_t> tiles;      // 4 words a tile
    std::vector<uint64_t> saddr;
    std::vector<uint32_t> steps;
    std::vector<uint32_t> part_of, lane_of;
    std::vector<double> body;         // per part: bundles a trip at its width
    std::vector<uint8_t> blob;        // the tables, when not staged through pinned memory
    uint32_t waves = 0, region_rows = 0;
    uint64_t blocks = 0;
    double makespan = 0, longest = 0;   // the planner's estimate, in bundles of one wave
    bool launched = false;
};

// Pinned host memory through which multi_launch uploads its tables, so the
// upload does not wait on the stream; kept by the caller until the launch has
// passed (then reusable).
struct PinnedStage {
    uint8_t *p = nullptr;
    size_t cap = 0;
    std::vector<uint8_t *> retired;   // outgrown buffers, freed with the stage
    PinnedStage() = default;
    PinnedStage(const PinnedStage &) = delete;
    PinnedStage &operator=(const PinnedStage &) = delete;
    PinnedStage(PinnedStage &&o) noexcept : p(o.p), cap(o.cap), retired(std::move(o.retired))
    {
        o.p = nullptr;
        o.cap = 0;
    }
    ~PinnedStage();
};

// The planner's knobs (0: its own choice).
struct MultiTune {
    uint32_t waves_per_group = 0;     // 1, 2 or 4
    uint32_t min_waves_per_cu = 0;    // plan for exactly this many waves a CU's LDS holds
    double alone = 0;                 // the issue fraction of a wave alone on a SIMD
    uint32_t max_waves_per_cu = 0;    // resident waves a CU the launch may take (room
                                      // left for other streams' kernels)
};

// Plan (widths, LDS region, grid) and enqueue one k_seq_multi launch on
// ex's stream; dlflags (device, one byte a position: at most the parts'
// lanes) or NULL.  Every part's program must be wide-eligible and of one
// format.  The tables go up through `pin` when given, else from pageable
// memory; with `pin` the caller keeps launches in flight, so arenas grow
// without waiting (arena_grow).  Does not wait.
cft_status multi_launch(cft_device *ex, std::vector<MultiPart> &parts, uint8_t *dlflags,
                        const MultiTune &tune, MultiRun &run, const char *who,
                        PinnedStage *pin = nullptr);

// Release a program's captured graph (the device current).
void seq_drop_graph(cft_program *p);

// Waves a workgroup and whether the code goes into LDS, for this program on
// this device (the largest workgroup up to waves_pref that fits).
struct Shape {
    int waves;
    bool code_in_lds;
};
Shape shape_for(const cft_program *p, int waves_pref);

// seq_check_bank (host/src/program.c:1640-1680)
cft_status seq_check_bank(const cft_program *p, const void *bank, size_t bank_bytes,
                          const char *who);

// Scratch slots a lane's run clears and the pool holds.
uint32_t seq_scratch_reach(const cft_program *p);
