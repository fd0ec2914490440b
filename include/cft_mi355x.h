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
 * 64) at a scratch depth (0 for 256) within max_vregs registers an instance
 * (0: no limit short of 65535), writes up to `cap` 16-byte step records
 * (steps x width) to `records` and the schedule's shape to info[0..5] =
 * {steps, registers, homes, segment-REPEAT step, body steps, body ALU
 * operations}.  CFT_ERR_UNSUPPORTED when the program is not pure (deposits,
 * activity, indexed scratch or flag control) or does not fit. */
CFT_API cft_status cftx_program_wide(const void *image, size_t bytes, uint32_t scratch_depth,
                                     uint32_t width, uint32_t max_vregs, void *records,
                                     size_t cap, uint32_t info[6]);

/* ---- several programs in one launch ------------------------------------------
 *
 * Instances of several pure programs (no deposits, activity, indexed
 * scratch or flag control) of one format run as one persistent launch: each
 * part is one program's batch with cftx_program_run_batch's meaning (lane i
 * runs steps[i] trips of the segment loop as the same chain of segments, its
 * state updated in place), and each lane's results are those of the same
 * lane run alone.  A planner picks each program's lanes an instance (unless
 * the part fixes it) and packs instances of every program into one
 * longest-first tile queue.  No reference counterpart: the reference runs
 * one program a call. */
typedef struct cftx_batch_part {
    size_t          struct_size;
    cft_program    *prog;
    const void     *bank;  size_t bank_bytes;   /* as cftx_program_run_batch */
    void           *state; size_t state_bytes;  /* n x slots elements, in place */
    const uint32_t *steps; size_t n;
    uint8_t        *lane_flags;          /* NULL or n bytes of host memory */
    uint32_t        lanes_per_instance;  /* 0: the planner's; else 1, 2, 4 ... 64 */
    uint32_t        width_used;          /* out: lanes an instance took */
} cftx_batch_part;

/* The planner's knobs, for an autotuner that sweeps them on a sample (0
 * everywhere, or a NULL pointer, leaves every choice to the planner):
 * waves_per_group 1, 2 or 4; waves_per_cu 1..12 plans for exactly that many
 * waves sharing a CU's LDS; alone in (0, 1] is the share of a SIMD's issue
 * rate the planner takes a wave alone on it to get. */
typedef struct cftx_multi_opts {
    size_t   struct_size;
    uint32_t waves_per_group;
    uint32_t waves_per_cu;
    double   alone;
} cftx_multi_opts;

/* flags_out: the OR of every lane's flags; bus_out: their STATUS bits. */
CFT_API cft_status cftx_program_run_multi(cft_device *dev, cftx_batch_part *parts,
                                          size_t nparts, const cftx_multi_opts *opts,
                                          uint32_t *flags_out, uint32_t *bus_out);

/* The plan of the handle's last multi-program launch: out[0..5] = {waves a
 * workgroup, workgroups, LDS rows a wave, tiles, workgroups a CU holds, the
 * planner's makespan estimate in bundles}. */
CFT_API cft_status cftx_multi_last_plan(cft_device *dev, uint64_t out[6]);

/* Build a pure program's wide schedules ahead of its first batched run: bit
 * k of `widths` asks for 2^k lanes an instance (k <= 6); *ready receives the
 * widths that schedule within the device's LDS.  CFT_ERR_UNSUPPORTED for a
 * program that is not pure.  A run builds what it needs on first use; this
 * takes that time out of a latency-sensitive loop. */
CFT_API cft_status cftx_program_prepare(cft_program *prog, uint32_t widths, uint32_t *ready);

/* ---- serving ---------------------------------------------------------------
 *
 * A request loop on one GPU with continuous batching.  Jobs submitted from
 * any thread are admitted while others run, by the device memory they need,
 * onto worker streams of the same GPU: with two or more workers, half of them
 * (rounded down) take elementwise jobs and the others program jobs; one
 * worker takes both, oldest job first.
 *  - Elementwise jobs waiting when an elementwise worker comes free go to it
 *    together, launched back to back.
 *  - Program jobs of pure programs join a program worker's running set (the
 *    worker already running jobs of the same program, else the least loaded
 *    one).  The
 *    worker runs the set in slices: one multi-program launch
 *    (cftx_program_run_multi's runner) in which every lane of every admitted
 *    job, whatever its program, advances by at most a slice's trips - a
 *    whole number of its segments, sized to the slice time - and new jobs
 *    join between slices.  A job completes when its last lane has run its
 *    steps.  Program jobs of other programs run whole, one batch a program.
 * Operands, outputs and states are windows of resident buffers (cft_alloc)
 * of the device the server was opened on, made current on the device
 * (cft_buffer_to_device) before submission, or windows of device memory of
 * that GPU the caller allocated itself (aligned to min(element size, 16)
 * bytes, inside one allocation); states are advanced in place and results
 * stay on the device.  An impure program job (one the wide runner does not
 * take) runs whole, alone when it has more than 2^22 lanes.  Each job's result is the bits and flags the
 * same job gives alone.  No reference counterpart (host/tools/cft-serve.c
 * answers one request at a time). */
#define CFTX_JOB_ELEM    0   /* d = op(a, b, c): cft_run's opcodes */
#define CFTX_JOB_DIV     1   /* d = a / b */
#define CFTX_JOB_SQRT    2   /* d = sqrt(a) */
#define CFTX_JOB_PROGRAM 3   /* a batched program run (cftx_program_run_batch) */

typedef struct cftx_server cftx_server;

typedef struct cftx_job {
    size_t          struct_size;
    int             kind;
    int             op, fmt, rnd;        /* elementwise */
    const void     *a, *b, *c;           /* device windows, n elements (NULL: unread) */
    void           *d;
    size_t          n;                   /* elements, or lanes of a program job */
    cft_program    *prog;                /* program jobs: loaded on the server's device */
    const void     *bank;  size_t bank_bytes;   /* host bytes, copied at submission */
    void           *state; size_t state_bytes;  /* a device window, n x slots elements */
    const uint32_t *steps;               /* n entries, copied at submission */
} cftx_job;

typedef struct cftx_job_result {
    size_t     struct_size;
    cft_status status;
    uint32_t   flags;        /* the job's exception flags */
    uint32_t   bus;          /* program jobs: the STATUS bits of its lanes */
    double     submit_s;     /* server clock (seconds) at submission */
    double     start_s;      /* at its first launch */
    double     done_s;       /* at completion, as the scheduler saw it */
    uint32_t   batch_jobs;   /* most jobs in a launch it ran in */
} cftx_job_result;

typedef struct cftx_server_stats {
    size_t   struct_size;
    uint64_t jobs, batches, elem_batches, program_batches;
    uint64_t max_batch_jobs;
    uint64_t peak_bytes;     /* most device memory admitted at once */
    uint64_t slices;         /* program slices launched (within program_batches) */
    uint64_t max_slice_lanes;
    double   slice_host_s;   /* scheduler time spent forming and launching slices */
} cftx_server_stats;

/* workers: streams (1..8, 0 for 4); hbm_budget: device bytes jobs may hold
 * at once beyond their resident windows (0 for 8 GiB). */
CFT_API cft_status cftx_server_open(cft_device *dev, int workers, size_t hbm_budget,
                                    cftx_server **out);
CFT_API cft_status cftx_server_submit(cftx_server *srv, const cftx_job *job,
                                      uint64_t *ticket);
/* Blocks until the job is done; each ticket is waited for once. */
CFT_API cft_status cftx_server_wait(cftx_server *srv, uint64_t ticket, cftx_job_result *out);
CFT_API double     cftx_server_clock(cftx_server *srv);
CFT_API cft_status cftx_server_stats_get(cftx_server *srv, cftx_server_stats *out);
/* Jobs a launch may take (1 serves every job alone, the plain form: one job
 * a launch, a program job run whole; 0 restores the default, 256 elementwise
 * jobs a batch and no limit on a program slice).  Applies to launches formed
 * after the call. */
CFT_API cft_status cftx_server_set_batching(cftx_server *srv, uint32_t max_jobs);
/* The time a program slice aims at, in milliseconds (default 5; 0 runs each
 * admitted job to its end in its first slice). */
CFT_API cft_status cftx_server_set_slice(cftx_server *srv, double ms);
/* Waits for every submitted job, then stops the workers. */
CFT_API void       cftx_server_close(cftx_server *srv);

/* The device address of a resident window (for tools that pair libcft with
 * other device code); NULL when p is not inside one of dev's buffers. */
CFT_API void *cftx_buffer_device_ptr(cft_device *dev, const void *p, size_t bytes);

/* cft_reduce (whole array) over operands already in device memory, with no
 * allocation at all.  No counterpart in the reference: it is for callers
 * that cannot allocate - code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  a, b (CFT_DOT only) and d are device
 * addresses of the handle's GPU (cftx_buffer_device_ptr, or the caller's own
 * device memory), aligned to min(element size, 16) bytes and each within one
 * allocation; the result is written to d on the device.  Arguments, results
 * and flags are cft_reduce's.  The tree's partial results live in a scratch
 * array of 8.5 MiB of device memory a GPU that the library holds until the
 * process ends, so n is at most 2^22 binary256, 2^23 binary128, 2^24 binary64
 * or 2^25 binary32 elements (CFT_ERR_INVALID_ARGUMENT beyond).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (two handles,
 * two threads) race, and their results are undefined. */
CFT_API cft_status cftx_reduce_noalloc(cft_device *dev, cft_op op, cft_format fmt,
                                       cft_round rnd, const void *a, const void *b, void *d,
                                       size_t n, uint32_t *flags_out, uint32_t *bus_out);

#ifdef __cplusplus
}
#endif

#endif /* CFT_MI355X_H */
