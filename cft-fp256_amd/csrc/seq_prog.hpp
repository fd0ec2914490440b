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
    // wide schedules by width, built on first use
    std::map<int, WideDev> wide;
    int wide_eligible = -1;          // -1 not yet known
};

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
