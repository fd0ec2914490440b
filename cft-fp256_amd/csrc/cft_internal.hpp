// Host side of the C ABI: the device handle, resident buffers, the error
// slot and the staging helpers every entry point shares.
#pragma once

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/cft.h"
#include "../../include/cft_mi355x.h"

// The library's own limits, published through cft_get_caps.
namespace cftl {
constexpr uint32_t kMaxDeposits = 1u << 20;     // deposit slots a lane
constexpr uint32_t kMaxInsns = 1u << 20;        // instructions an image holds
constexpr uint32_t kMaxConsts = 512;            // constants an insn addresses
constexpr uint32_t kDefaultDepth = 256;         // scratch slots a lane
constexpr uint32_t kMaxDepth = 2048;
constexpr uint32_t kSeqFeatures =
    CFT_SEQ_FEAT_WIDE_CONST | CFT_SEQ_FEAT_REGS32 | CFT_SEQ_FEAT_BANK_PTR | CFT_SEQ_FEAT_KX9 |
    CFT_ALU_EXT_IMUL | CFT_SEQ_FEAT_SCRATCH | CFT_SEQ_FEAT_SCRATCH_IO |
    CFT_SEQ_FEAT_SCRATCH_STRICT | CFT_SEQ_FEAT_SCALAR | CFT_SEQ_FEAT_LANE_MASK |
    CFT_SEQ_FEAT_INDEXED | CFT_FEAT_REDUCE_SEG |
    CFT_SEQ_FEAT_SCRATCH_STEP | CFT_SEQ_FEAT_LANE_FLAGS | CFT_SEQ_FEAT_FLAG_CONTROL;
}

// Staging roles: each operand a call binds has its own growable arena.
enum Role {
    ROLE_A, ROLE_B, ROLE_C, ROLE_D,                 // streams and the result/deposits
    ROLE_IA, ROLE_IB, ROLE_IC, ROLE_ISI,            // index tables
    ROLE_BANK, ROLE_SIN, ROLE_SOUT, ROLE_COUNTS, ROLE_MASK, ROLE_LFLAGS,
    ROLE_POOL, ROLE_TRIPS, ROLE_ORDER, ROLE_REGS,
    ROLE_MPROGS,                                    // the multi-program runner's tables
    ROLE_RED0, ROLE_RED1, ROLE_RWORDS,              // a reduction's partials and flag words
    ROLE_COUNT
};

// One growable device allocation.
struct DevArena {
    void *p = nullptr;
    size_t cap = 0;
};

struct cft_buffer {
    cft_device *dev = nullptr;
    void *host = nullptr;          // the mirror the caller holds (pinned)
    size_t bytes = 0;
    void *devp = nullptr;          // the device copy
    bool host_valid = true;        // mirror holds the current bytes
    bool dev_valid = false;        // device copy holds the current bytes
    uint64_t resident_binds = 0, staged_binds = 0;
    char why[112] = {0};
};

struct cft_device {
    int hip_index = 0;
    hipStream_t stream = nullptr;
    uint32_t sticky = 0;           // 754 7.1 status word
    bool poisoned = false;
    uint32_t scratch_depth = cftl::kDefaultDepth;
    int num_cus = 256;
    DevArena stage[ROLE_COUNT];    // per operand role
    size_t lds_max = 65536;        // LDS bytes a workgroup may take
    uint32_t *dwords = nullptr;    // device: [0] flags, [1] status
    uint32_t *hwords = nullptr;    // pinned mirror of dwords
    uint32_t *hwords_dev = nullptr;  // hwords' device address
    // the elementwise calls' flag slots (kernels.h), zero between calls
    uint32_t *dslots = nullptr;
    std::vector<cft_buffer *> bufs;
    // cftx_kernel_timer: events around each compute kernel on `stream`
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double timed_ms = 0.0;
    uint64_t timed_launches = 0;
    uint64_t batch_launches = 0;   // kernel launches the last batched run made
    uint32_t batch_width = 0;      // lanes an instance took in it
    // arenas replaced while launches that read them were in flight, freed
    // with the handle (freeing device memory waits for the whole device)
    std::vector<void *> retired;
    // the last multi-program launch's plan: waves a group, groups, LDS rows a
    // wave, tiles, groups a CU holds, planned makespan (bundles)
    uint64_t multi_plan[6] = {0, 0, 0, 0, 0, 0};
#ifdef CFT_PLANTS
    int plant_fail_launch = 0;     // CFT_PLANT=fail-launch: the next launch fails
#endif
};

// Before a handle's compute launch.  hipSuccess, except in the negative-
// control build (-DCFT_PLANTS, lib/libcft_plants.so), where a handle opened
// with CFT_PLANT=fail-launch has its first launch refused here, before
// anything is launched, so the failure path runs with no kernel involved.
hipError_t plant_launch(cft_device *dev);

// Bracket a compute-kernel launch with the handle's timer events (no-ops
// unless cftx_kernel_timer enabled them); timer_end adds the elapsed time
// once the stream has been synchronised.  timer_begin calls plant_launch.
hipError_t timer_begin(cft_device *dev);
hipError_t timer_end(cft_device *dev);
void timer_collect(cft_device *dev);

// ---- the error slot (process-global, as the reference's) ----------------
void cft_clear_error();
void cft_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Emit a call's flags: OR into the device's word, write flags_out.
void cft_emit_flags(cft_device *dev, uint32_t acc, uint32_t *flags_out);

// ---- device plumbing -------------------------------------------------------

// Make dev's GPU current.  CFT_ERR_INTERNAL (with a sentence) on a poisoned
// handle.
cft_status dev_enter(cft_device *dev, const char *who);

// A HIP failure after work may have started: poison the handle.
cft_status dev_fail(cft_device *dev, hipError_t e, const char *who, const char *what);

// Grow arena to at least bytes.  CFT_ERR_OUT_OF_MEMORY before anything ran.
cft_status arena_reserve(cft_device *dev, DevArena &a, size_t bytes, const char *who);

// The same without waiting: a larger arena replaces the old one, which is
// kept (retired) until the handle closes, so launches still queued on the
// stream keep reading valid memory.  Grows at least twofold.
cft_status arena_grow(cft_device *dev, DevArena &a, size_t bytes, const char *who);

// Operand binding.  An input window [p, p+bytes) that lies inside a resident
// buffer is served from its device copy (refreshed first if stale); any other
// window is copied into the role's staging arena.  *dptr receives the device
// address.
cft_status bind_in(cft_device *dev, int role, const void *p, size_t bytes, const void **dptr,
                   const char *who);
// Output window: a buffer's device copy (made current first), or the role's
// arena; *buf receives the buffer or NULL (then the caller copies back).
cft_status bind_out(cft_device *dev, int role, void *p, size_t bytes, void **dptr,
                    cft_buffer **buf, const char *who);

// The buffer containing [p, p+bytes), or NULL.
cft_buffer *find_buffer(cft_device *dev, const void *p, size_t bytes);

// Before the host reads [p, p+bytes): a resident buffer whose start p points
// into has its mirror brought up to date from the device if a run wrote it.
cft_status host_in(cft_device *dev, const void *p, const char *who);

// Another handle on GPU hip_index at scratch depth `depth` (its own stream
// and staging arenas), for the serving runtime's workers; own_queue asks for
// a stream on a hardware queue no other stream shares.
cft_status device_open(int hip_index, uint32_t depth, cft_device **out, bool own_queue = false);

// Kernel grid for n lanes of one element each.
int grid_for(const cft_device *dev, size_t n);

// A window [p, p + bytes) of raw device memory (not a resident buffer):
// aligned to `align` bytes, device memory of dev's GPU, and inside one
// allocation.  CFT_ERR_INVALID_ARGUMENT with a sentence naming `what`
// otherwise.
cft_status device_window(const char *who, cft_device *dev, const char *what, const void *p,
                         size_t bytes, size_t align);

// cft_reduce's argument checks (host/src/device.c:2759-2862) in their order:
// dev, format, rounding attribute, a reduction opcode, d, and for n > 0 the
// operands the opcode reads and a byte count a size_t holds.
cft_status reduce_check(const char *who, cft_device *dev, int op, int fmt, int rnd,
                        const void *a, const void *b, void *d, size_t n);

// Which of a, b, c (bits 0, 1, 2) elementwise opcode op reads
// (host/src/softfloat.c:822-848); 0 for an unassigned opcode.
unsigned cft_op_reads(int op);
