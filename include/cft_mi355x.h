/* cft_mi355x.h - extensions of the MI355X engine that have no counterpart
 * in the reference ABI (host/include/cft.h).  Callers that only use the
 * reference's entry points never need this header.
 */
#ifndef CFT_MI355X_H
#define CFT_MI355X_H

#include "cft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel timing on the handle's own stream.  While enabled, every
 * compute-kernel launch the handle makes is bracketed by HIP events on the
 * stream it is launched on, and the elapsed times are summed.  For the
 * benchmark's roofline figure (kernel time, not call time).  No reference
 * counterpart: the reference times calls, not kernels. */
CFT_API cft_status cftx_kernel_timer(cft_device *dev, int enable);
CFT_API cft_status cftx_kernel_timer_read(cft_device *dev, double *total_ms,
                                          uint64_t *launches);

/* The loader's checks without a device: parse and validate a program image
 * against a scratch depth (0 for the default, 256), returning the status
 * cft_program_load would return on a handle opened at that depth, and its
 * cft_program_info where info is non-NULL and accepted.  For tools that
 * vet images offline, and for testing the loader on a host with no GPU.
 * No reference counterpart: the reference's loader needs a device handle. */
CFT_API cft_status cftx_program_check(const void *image, size_t bytes,
                                      uint32_t scratch_depth,
                                      cft_program_info *info);

/* ---- batched program runs ------------------------------------------------
 *
 * Many instances of one program, each of uneven length, in one call: lane i
 * runs steps[i] trips of the program's first top-level REPEAT (its segment
 * loop) as a chain of segments of at most that REPEAT's own count - exactly
 * the chain of cft_program_run_ex calls that carries a state from each
 * segment's scratch-out block into the next one's scratch-in block, with the
 * last segment's REPEAT count the remainder.  The program must declare
 * scratch I/O with n_scratch_in == n_scratch_out (the state) and deposit
 * nothing.  No reference counterpart: the reference chains segments one run
 * at a time (host/tools/orbits.c, segrun.c).
 *
 * Modes: CFTX_BATCH_PERSISTENT runs the whole batch in one launch of
 * persistent workgroups that take tiles of lanes (longest first) from a
 * counter; CFTX_BATCH_PER_OP is the plain form for comparison - registers
 * and scratch in HBM and one kernel launch per instruction, the host
 * walking the program.  Both give the same bits. */
#define CFTX_BATCH_PERSISTENT 0
#define CFTX_BATCH_PER_OP     1

typedef struct cftx_batch_args {
    size_t          struct_size;
    const void     *bank;        size_t bank_bytes;    /* BANK_EXT programs */
    void           *state;       size_t state_bytes;   /* n x n_in elements, in place */
    const uint32_t *steps;                             /* n entries */
    size_t          n;
    const void     *a, *b, *c;   /* r0..r2 of lane i: element i, or +0 where NULL */
    uint32_t       *flags_out;
    uint32_t       *bus_out;
    uint8_t        *lane_flags;  size_t lane_flags_bytes;  /* n bytes: OR over a lane's segments */
    int             mode;
    uint32_t        lanes_per_wave;   /* 1..64 lanes a wavefront takes; 0 chooses */
    uint32_t        waves_per_group;  /* 1, 2 or 4; 0 chooses */
    int             graph;            /* launch through a captured HIP graph (persistent) */
    uint32_t        lanes_per_instance; /* persistent form: 1 interprets an instance on one
                                           lane; 2..64 (a power of two) spreads it over that
                                           many lanes (a pure program only); 0 chooses */
} cftx_batch_args;

CFT_API cft_status cftx_program_run_batch(cft_program *prog, const cftx_batch_args *args);

/* Kernel launches the handle's last cftx_program_run_batch made (2 for the
 * persistent form: the counter reset and the runner). */
CFT_API uint64_t cftx_batch_launches(cft_device *dev);

/* Lanes an instance took in the handle's last persistent batched run. */
CFT_API uint32_t cftx_batch_width(cft_device *dev);

/* The wide schedule of an image without a device, for tools and tests:
 * schedules the program for `width` lanes an instance (a power of two up to
 * 64) at a scratch depth (0 for 256), writes up to `cap` 16-byte step
 * records (steps x width) to `records` and the schedule's shape to info[0..5]
 * = {steps, registers, homes, segment-REPEAT step, body steps, body ALU
 * operations}.  CFT_ERR_UNSUPPORTED when the program is not pure (deposits,
 * activity, indexed scratch or flag control) or does not fit. */
CFT_API cft_status cftx_program_wide(const void *image, size_t bytes, uint32_t scratch_depth,
                                     uint32_t width, void *records, size_t cap,
                                     uint32_t info[6]);

#ifdef __cplusplus
}
#endif

#endif /* CFT_MI355X_H */
