// Wide execution of a sequencer program: one program instance spread over W
// lanes of a wavefront.
//
// The loader-side scheduler (seq_wide.cpp) turns the decoded instruction
// stream of a "pure" program (ALU instructions, static scratch accesses,
// loops; no deposits, activity, indexed scratch or flag control) into a
// sequence of steps.  A step is either a bundle of up to W independent
// operations - lane j of an instance's group executes operation j - or a
// control step (REPEAT / ENDREP / HALT), uniform over the wave.
//
// Within each straight-line block the scheduler renames every value (so
// register reuse in the source creates no false dependence), forwards
// stores to later loads of the same slot, loads each slot's entry value
// once, keeps only the stores and register values that are live after the
// block (liveness over the loop structure, the scratch-out block live at the
// end), and list-schedules by critical path.  Every ALU instruction that can
// raise a flag is kept, since its flags are observable; ABS and NEG, which
// change a sign bit and raise none, become sign modifiers on their readers'
// operands (dropped when nothing else reads their result), which takes them
// off the dependence chain.  Each result is the same correctly rounded
// operation on the same operands, so the values, and the OR of the flags,
// are those of the source program in source order.
//
// Registers of an instance: homes [0, nregs) hold the program's (renamed)
// registers between blocks; temporaries [nregs, nvregs) hold values inside
// a block.
#pragma once

#include <stdint.h>

#include <vector>

// a step's record for one lane position: 16 bytes
struct WideRec {
    uint32_t x;   // [7:0] op, [10:8] kind, [13:11] rnd, [14] ka, [15] kb, [16] kc,
                  // [18:17] [20:19] [22:21] sign modifiers of a, b, c (bit 0 clear
                  // the sign, bit 1 flip it, in that order: a fused ABS / NEG)
    uint32_t y;   // [15:0] a, [31:16] b
    uint32_t z;   // [15:0] c, [31:16] dst
    uint32_t w;   // control: REPEAT trips; unused otherwise
};

enum {
    WK_ALU = 0,     // dst <- op(a, b, c), a/b/c registers or constants
    WK_LOAD = 1,    // dst <- scratch[a]
    WK_STORE = 2,   // scratch[dst] <- register a
    WK_MOVE = 3,    // dst <- register a
    WK_NOP = 4,
    WK_CTRL = 5     // op: SEQ_REPEAT (w trips, y the matching ENDREP step),
                    // SEQ_ENDREP (y the first body step), SEQ_HALT
};

#define WR_OP(x)   ((x) & 0xFFu)
#define WR_KIND(x) (((x) >> 8) & 7u)
#define WR_RND(x)  (((x) >> 11) & 7u)
#define WR_KA(x)   (((x) >> 14) & 1u)
#define WR_KB(x)   (((x) >> 15) & 1u)
#define WR_KC(x)   (((x) >> 16) & 1u)
#define WR_MODS(x) (((x) >> 17) & 0x3Fu)
#define WR_A(y)    ((y) & 0xFFFFu)
#define WR_B(y)    ((y) >> 16)
#define WR_C(z)    ((z) & 0xFFFFu)
#define WR_DST(z)  ((z) >> 16)

struct WideProgram {
    int width = 0;                  // W: lanes an instance takes
    uint32_t nsteps = 0;
    uint32_t nregs = 0;             // homes
    uint32_t nvregs = 0;            // homes + temporaries
    uint32_t top_step = 0xFFFFFFFFu;  // the segment REPEAT's step
    uint32_t body_steps = 0;        // steps of the segment loop's body
    uint32_t body_alu = 0;          // ALU operations in it
    std::vector<WideRec> recs;      // nsteps x width
};

struct cft_program;

// Whether the program can run wide (pure: see above).
bool seq_wide_eligible(const cft_program *p);

// The scheduler's lookahead windows, widest first: a narrower window keeps
// fewer values live (fewer registers) at the cost of some parallelism.
extern const int kWideWindows[];
extern const int kWideWindowCount;

// Schedule for `width` lanes an instance at one window; false when the
// program does not schedule.
bool seq_wide_compile_at(const cft_program *p, int width, int window, WideProgram *out);

// Schedule for `width` lanes an instance at the widest window within
// `max_vregs` registers; false when none is.
bool seq_wide_compile(const cft_program *p, int width, uint32_t max_vregs, WideProgram *out);
