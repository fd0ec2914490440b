// The sequencer's pre-decoded instruction record, shared by the host loader
// (cft_prog.cpp) and the interpreter kernel (k_seq.hip).
//
// The image's 64-bit words (op[7:0], rd[11:8], ra/rb/rc, rnd[26:24],
// ka/kb/kc/kx[30:27], ctrl[31], imm[63:32] with the fifth register bits in
// imm[27:24] and the ninth constant-index bits in imm[30:28]; host/src/
// seq_exec.h:175-248) are decoded once at load into this form, so the
// interpreter reads resolved register numbers and constant indices and never
// scans for a loop's end.
#pragma once

#include <stdint.h>

struct SeqInsn {
    uint32_t w;     // [7:0] op, [8] ctrl, [11:9] rnd, [12] ka, [13] kb, [14] kc,
                    // [20:16] rd
    uint32_t src;   // [8:0] a, [17:9] b, [26:18] c: register 0..31, or
                    // constant index 0..511 where the k bit is set
    uint32_t imm;   // REPEAT trip count; STL/LDL slot; STX/LDX post-step (int32)
    uint32_t aux;   // REPEAT: index of its ENDREP; ENDREP: index of the body
};

// control codes (host/src/seq_exec.h:63-72)
enum {
    SEQ_HALT = 0, SEQ_REPEAT, SEQ_ENDREP, SEQ_DEPOSIT, SEQ_SETACT, SEQ_ACTALL, SEQ_STL,
    SEQ_LDL, SEQ_STX, SEQ_LDX, SEQ_AUGADD, SEQ_AUGERR, SEQ_QUIET, SEQ_ENDQUIET, SEQ_RAISE
};

// run status (bus_out) and lane-byte bits: cft.h's CFT_STATUS_* / CFT_LANE_*
#define SEQ_ST_DEPOSIT_OVERFLOW  0x10u
#define SEQ_ST_SCRATCH_RANGE     0x20u
#define SEQ_ST_MARKED            0x40u
#define SEQ_LF_DEPOSIT_OVERFLOW  0x20u
#define SEQ_LF_SCRATCH_RANGE     0x40u
#define SEQ_LF_MARKED            0x80u

#define SEQ_W_OP(w)    ((w) & 0xFFu)
#define SEQ_W_CTRL(w)  (((w) >> 8) & 1u)
#define SEQ_W_RND(w)   (((w) >> 9) & 7u)
#define SEQ_W_KA(w)    (((w) >> 12) & 1u)
#define SEQ_W_KB(w)    (((w) >> 13) & 1u)
#define SEQ_W_KC(w)    (((w) >> 14) & 1u)
#define SEQ_W_RD(w)    (((w) >> 16) & 31u)
#define SEQ_SRC_A(s)   ((s) & 511u)
#define SEQ_SRC_B(s)   (((s) >> 9) & 511u)
#define SEQ_SRC_C(s)   (((s) >> 18) & 511u)

// Everything one interpreter launch reads and writes.  Device pointers.
struct SeqArgs {
    const SeqInsn *code;
    uint32_t n_insns;
    uint32_t n_consts;
    const uint32_t *bank;             // n_consts x NL limbs
    const uint32_t *a, *b, *c;        // the streams, element i at i*NL (b, c may be NULL)
    const uint32_t *idx_a, *idx_b, *idx_c;
    const uint32_t *sin;              // scratch-in, lane-major (or its pool)
    const uint32_t *idx_si;
    uint32_t n_in, n_out;
    uint32_t *sout;                   // scratch-out, lane-major
    uint32_t *pool;                   // scratch: [slot][limb][lane], lane fastest
    uint64_t npool;                   // lanes the pool is laid out for
    uint32_t reach;                   // slots zeroed before the run
    uint32_t depth;                   // power of two
    int log2depth;
    int strict;
    uint32_t *dep;                    // deposits, n x max_dep elements
    uint32_t max_dep;
    uint32_t *counts;
    const uint8_t *mask;              // lane mask bits, or NULL
    uint8_t *lflags;                  // lane flag bytes, or NULL
    uint64_t n;                       // lanes in this launch
    uint64_t lane0;                   // global index of this launch's lane 0
    uint32_t *words;                  // [0] run flags, [1] run status (ORed)
    int code_in_lds;                  // stage the code in LDS (it fits)
    // per-lane trip count of the program's first top-level REPEAT, or NULL
    // for the image's own count
    const uint32_t *trips;
};

// The batched runner's extra arguments (k_seq_batch): lanes are handed out
// in tiles of `lpw` to persistent waves through an atomic counter, each lane
// runs its `steps` as a chain of segments of at most `seg` trips of the
// segment REPEAT, its state (n_in == n_out slots) read before and written
// after every segment.
struct SeqBatch {
    const uint32_t *order;            // lane ids, tile-major (longest first)
    const uint32_t *steps;            // per lane id
    uint32_t *state;                  // per lane id: n_in slots, lane-major
    uint32_t *queue;                  // tile counter, 0 at launch
    uint8_t *lflags;                  // per lane id, OR over its segments, or NULL
    uint64_t n;                       // lanes
    uint32_t ntiles;
    uint32_t lpw;                     // lanes a wave takes (1..64)
    uint32_t seg;                     // trips a segment
    uint32_t hi_steps;                // a wave whose tile holds a lane with at
                                      // least this many trips issues at raised
                                      // priority (0: every wave at the default)
};

// The wide batched runner's program (seq_wide.hpp): `nsteps` steps of
// `width` 16-byte records, instance registers [0, nvregs).
struct SeqWide {
    const uint4 *recs;
    uint32_t nsteps;
    uint32_t width;                   // lanes an instance takes (power of two)
    uint32_t log2w;
    uint32_t nvregs;
    uint32_t top_step;                // the segment REPEAT's step
    int recs_in_lds;
};

// The multi-program runner (k_seq_multi): instances of several wide
// programs of one format in one launch.  A program's entry (64 bytes):
struct SeqMProg {
    const uint4 *recs;                // nsteps x width records
    const uint32_t *bank;             // its constants, n_consts elements
    uint32_t nsteps;
    uint32_t log2w;                   // lanes an instance takes: 1 << log2w
    uint32_t nvregs;
    uint32_t top_step;
    uint32_t n_consts;
    uint32_t n_in;                    // == n_out: the state's slots
    uint32_t reach;                   // scratch slots cleared a segment
    uint32_t seg;                     // trips a segment
    uint32_t pad[4];
};

// Tiles are (program, first position, instances) triples handed out longest
// first; a position names an instance's state and trips.
struct SeqMulti {
    const SeqMProg *progs;
    const uint4 *tiles;               // x program, y first position, z count
    const uint64_t *saddr;            // per position: its state's device address
    const uint32_t *steps;            // per position: trips
    uint8_t *lflags;                  // per position: flag byte, or NULL
    uint32_t *queue;                  // tile counter, 0 at launch
    uint32_t ntiles;
    uint32_t region_rows;             // LDS rows a wave takes (constants + registers)
};
