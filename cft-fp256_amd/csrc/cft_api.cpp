// The C ABI's general entry points: names, the error slot, device handles,
// capabilities, the 754 status word and resident buffers.
//
// Behaviour mirrors the reference's device.c (open/caps/flags/buffers,
// host/src/device.c:360-420, 653-760, 3700-3800) and its GPU backend's open
// grammar and failure discipline (host/src/backend_hip.cpp:1036-1096,
// 2760-2830): argument errors are CFT_ERR_INVALID_ARGUMENT with a sentence,
// a HIP failure after work may have started poisons the handle, an
// allocation refused before anything ran is CFT_ERR_OUT_OF_MEMORY and the
// handle survives.
#include "cft_internal.hpp"
#include "kernels.h"

#include <limits.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include <algorithm>
#include <new>

// ---- error slot -----------------------------------------------------------

static char g_err[512];

void cft_clear_error() { g_err[0] = '\0'; }

void cft_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" CFT_API const char *cft_last_error(void) { return g_err; }

extern "C" CFT_API uint32_t cft_abi_version(void)
{
    return ((uint32_t)CFT_ABI_VERSION_MAJOR << 16) | (uint32_t)CFT_ABI_VERSION_MINOR;
}

// host/src/device.c:368-388
extern "C" CFT_API const char *cft_strerror(cft_status s)
{
    switch (s) {
    case CFT_OK: return "ok";
    case CFT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CFT_ERR_UNSUPPORTED: return "operation or format not available on this device";
    case CFT_ERR_NO_DEVICE: return "no such device";
    case CFT_ERR_ARTIFACT: return "artifact missing, unreadable, or not a tile";
    case CFT_ERR_BUS_FAULT: return "memory system fault: the output is not valid";
    case CFT_ERR_OUT_OF_MEMORY: return "out of memory";
    case CFT_ERR_TIMEOUT: return "timed out";
    case CFT_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

extern "C" CFT_API size_t cft_format_size(cft_format f)
{
    static const size_t sz[4] = {4, 8, 16, 32};
    return ((int)f >= 0 && (int)f <= 3) ? sz[(int)f] : 0;
}

extern "C" CFT_API const char *cft_format_name(cft_format f)
{
    static const char *const nm[4] = {"fp32", "fp64", "fp128", "fp256"};
    return ((int)f >= 0 && (int)f <= 3) ? nm[(int)f] : "invalid";
}

// The lower-case names the reference publishes (host/src/device.c:401-425),
// which its Python binding audits its opcode numbers against.
extern "C" CFT_API const char *cft_op_name(cft_op op)
{
    static const char *const names[32] = {
        "fma", "add", "sub", "mul", "abs", "neg", "copysign", "min", "max", "minnum",
        "maxnum", "select", "cmplt", "cmple", "cmpeq", nullptr, "iand", "ior", "ixor", "iadd",
        "isub", "ishl", "ishr", "icmplt", "sum", "dot", "recip_seed", "rsqrt_seed", "sumsq",
        "sumabs", "imul", "maxall"};
    const int i = (int)op;
    return (i >= 0 && i < 32 && names[i]) ? names[i] : "reserved";
}

// ---- flags ----------------------------------------------------------------

void cft_emit_flags(cft_device *dev, uint32_t acc, uint32_t *flags_out)
{
    if (dev)
        dev->sticky |= acc & CFT_FLAGS_ALL;
    if (flags_out)
        *flags_out = acc;
}

extern "C" CFT_API void cft_lower_flags(cft_device *dev, uint32_t mask)
{
    if (dev) dev->sticky &= ~(mask & CFT_FLAGS_ALL);
}

extern "C" CFT_API void cft_raise_flags(cft_device *dev, uint32_t mask)
{
    if (dev) dev->sticky |= mask & CFT_FLAGS_ALL;
}

extern "C" CFT_API int cft_test_flags(cft_device *dev, uint32_t mask)
{
    return dev && (dev->sticky & mask & CFT_FLAGS_ALL) ? 1 : 0;
}

extern "C" CFT_API uint32_t cft_save_all_flags(cft_device *dev) { return dev ? dev->sticky : 0; }

extern "C" CFT_API void cft_restore_flags(cft_device *dev, uint32_t saved, uint32_t mask)
{
    if (!dev)
        return;
    mask &= CFT_FLAGS_ALL;
    dev->sticky = (dev->sticky & ~mask) | (saved & mask);
}

extern "C" CFT_API int cft_test_saved_flags(uint32_t saved, uint32_t mask)
{
    return (saved & mask & CFT_FLAGS_ALL) ? 1 : 0;
}

// ---- device plumbing --------------------------------------------------------

cft_status dev_enter(cft_device *dev, const char *who)
{
    if (dev->poisoned) {
        cft_set_error("%s: this handle saw a GPU failure after work had started, so "
                      "nothing it holds can be trusted; close it and open it again",
                      who);
        return CFT_ERR_INTERNAL;
    }
    const hipError_t e = hipSetDevice(dev->hip_index);
    if (e != hipSuccess) {
        cft_set_error("%s: selecting GPU %d failed: %s", who, dev->hip_index,
                      hipGetErrorString(e));
        (void)hipGetLastError();
        return CFT_ERR_NO_DEVICE;
    }
    return CFT_OK;
}

cft_status dev_fail(cft_device *dev, hipError_t e, const char *who, const char *what)
{
    dev->poisoned = true;
    cft_set_error("%s: %s failed on GPU %d: %s (%s); the handle is poisoned - close it "
                  "and open it again",
                  who, what, dev->hip_index, hipGetErrorString(e), hipGetErrorName(e));
    (void)hipGetLastError();
    return CFT_ERR_INTERNAL;
}

cft_status arena_grow(cft_device *dev, DevArena &a, size_t bytes, const char *who)
{
    if (bytes <= a.cap)
        return CFT_OK;
    size_t want = std::max<size_t>({bytes, a.cap * 2, (size_t)1 << 20});
    want = (want + 4095) & ~(size_t)4095;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        want = bytes;
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        cft_set_error("%s: %zu bytes of device memory could not be allocated on GPU %d", who,
                      bytes, dev->hip_index);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    if (a.p)
        dev->retired.push_back(a.p);
    a.p = p;
    a.cap = want;
    return CFT_OK;
}

cft_status arena_reserve(cft_device *dev, DevArena &a, size_t bytes, const char *who)
{
    if (bytes <= a.cap)
        return CFT_OK;
    // grow geometrically so a sweep of sizes settles quickly
    size_t want = std::max(bytes, a.cap + a.cap / 2);
    want = (want + 4095) & ~(size_t)4095;
    if (a.p) {
        // the stream may still be reading it
        const hipError_t e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "waiting for the stream");
        (void)hipFree(a.p);
        a.p = nullptr;
        a.cap = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMalloc(&p, bytes);               // exactly what is needed
        want = bytes;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        cft_set_error("%s: %zu bytes of device memory could not be allocated on GPU %d",
                      who, bytes, dev->hip_index);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    a.p = p;
    a.cap = want;
    return CFT_OK;
}

// One workgroup of 256 lanes for every 256 elements: with the flag slots,
// binary256 FMA/ADD/MUL/DIV over 2^24 elements run 7-12% faster this way
// than on a grid capped at 8 workgroups a CU (a grid-stride loop), and
// 2^20 binary128 elements take the same time either way.
int grid_for(const cft_device *dev, size_t n)
{
    (void)dev;
    const size_t want = (n + 255) / 256;
    return (int)std::max<size_t>(1, std::min<size_t>(want, (size_t)INT_MAX));
}

cft_status device_window(const char *who, cft_device *dev, const char *what, const void *p,
                        size_t bytes, size_t align)
{
    if ((uintptr_t)p % align) {
        cft_set_error("%s: %s (%p) is not aligned to %zu bytes", who, what, p, align);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess || at.type != hipMemoryTypeDevice || at.device != dev->hip_index) {
        (void)hipGetLastError();
        cft_set_error("%s: %s (%p) is not device memory of GPU %d", who, what, p,
                      dev->hip_index);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
    if (e != hipSuccess || (const uint8_t *)p + bytes > (const uint8_t *)base + size) {
        (void)hipGetLastError();
        cft_set_error("%s: %s's %zu bytes at %p run past the end of its allocation", who, what,
                      bytes, p);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return CFT_OK;
}

// host/src/softfloat.c:822-848: FMA and SELECT read all three, ADD/SUB a and
// c, the sign operations and the seeds a alone, every other assigned opcode
// a and b.
unsigned cft_op_reads(int op)
{
    const bool assigned = (op >= 0 && op <= 14) || (op >= 16 && op <= 31);
    switch (op) {
    case 0: case 11: return 7u;
    case 1: case 2: return 5u;
    case 4: case 5: case 26: case 27: return 1u;
    default: return assigned ? 3u : 0u;
    }
}

cft_buffer *find_buffer(cft_device *dev, const void *p, size_t bytes)
{
    const uint8_t *q = (const uint8_t *)p;
    for (cft_buffer *b : dev->bufs) {
        const uint8_t *h = (const uint8_t *)b->host;
        if (q >= h && q + bytes <= h + b->bytes && bytes <= b->bytes)
            return b;
    }
    return nullptr;
}

cft_status host_in(cft_device *dev, const void *p, const char *who)
{
    if (!p)
        return CFT_OK;
    const uint8_t *q = (const uint8_t *)p;
    for (cft_buffer *b : dev->bufs) {
        const uint8_t *h = (const uint8_t *)b->host;
        if (q < h || q >= h + b->bytes || b->host_valid)
            continue;
        cft_status st = dev_enter(dev, who);
        if (st != CFT_OK)
            return st;
        hipError_t e = hipMemcpyAsync(b->host, b->devp, b->bytes, hipMemcpyDeviceToHost,
                                      dev->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess)
            return dev_fail(dev, e, who, "bringing a buffer's mirror up to date");
        b->host_valid = true;
    }
    return CFT_OK;
}

// A buffer's device copy made current before a run binds it.
static cft_status buffer_fresh(cft_buffer *b, const char *who, bool *staged)
{
    *staged = false;
    if (b->dev_valid)
        return CFT_OK;
    const hipError_t e = hipMemcpyAsync(b->devp, b->host, b->bytes, hipMemcpyHostToDevice,
                                        b->dev->stream);
    if (e != hipSuccess)
        return dev_fail(b->dev, e, who, "refreshing a buffer's device copy");
    b->dev_valid = true;
    *staged = true;
    snprintf(b->why, sizeof b->why, "the device copy was stale: the mirror was written "
                                    "since the last transfer");
    return CFT_OK;
}

cft_status bind_in(cft_device *dev, int role, const void *p, size_t bytes, const void **dptr,
                   const char *who)
{
    if (!p) {
        *dptr = nullptr;
        return CFT_OK;
    }
    if (cft_buffer *b = find_buffer(dev, p, bytes)) {
        bool staged;
        const cft_status st = buffer_fresh(b, who, &staged);
        if (st != CFT_OK)
            return st;
        if (staged) b->staged_binds++; else b->resident_binds++;
        *dptr = (const uint8_t *)b->devp + ((const uint8_t *)p - (const uint8_t *)b->host);
        return CFT_OK;
    }
    cft_status st = arena_reserve(dev, dev->stage[role], bytes, who);
    if (st != CFT_OK)
        return st;
    const hipError_t e =
        hipMemcpyAsync(dev->stage[role].p, p, bytes, hipMemcpyHostToDevice, dev->stream);
    if (e != hipSuccess)
        return dev_fail(dev, e, who, "copying an operand to the GPU");
    *dptr = dev->stage[role].p;
    return CFT_OK;
}

cft_status bind_out(cft_device *dev, int role, void *p, size_t bytes, void **dptr,
                    cft_buffer **buf, const char *who)
{
    if (cft_buffer *b = find_buffer(dev, p, bytes)) {
        bool staged;
        const cft_status st = buffer_fresh(b, who, &staged);
        if (st != CFT_OK)
            return st;
        if (staged) b->staged_binds++; else b->resident_binds++;
        *dptr = (uint8_t *)b->devp + ((uint8_t *)p - (uint8_t *)b->host);
        *buf = b;
        return CFT_OK;
    }
    cft_status st = arena_reserve(dev, dev->stage[role], bytes, who);
    if (st != CFT_OK)
        return st;
    *dptr = dev->stage[role].p;
    *buf = nullptr;
    return CFT_OK;
}

// ---- kernel timer (include/cft_mi355x.h) -------------------------------------

hipError_t plant_launch(cft_device *dev)
{
#ifdef CFT_PLANTS
    if (dev->plant_fail_launch) {
        dev->plant_fail_launch = 0;
        return hipErrorLaunchFailure;
    }
#else
    (void)dev;
#endif
    return hipSuccess;
}

hipError_t timer_begin(cft_device *dev)
{
    const hipError_t e = plant_launch(dev);
    if (e != hipSuccess)
        return e;
    return dev->timing ? hipEventRecord(dev->ev0, dev->stream) : hipSuccess;
}

hipError_t timer_end(cft_device *dev)
{
    return dev->timing ? hipEventRecord(dev->ev1, dev->stream) : hipSuccess;
}

void timer_collect(cft_device *dev)
{
    if (!dev->timing)
        return;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, dev->ev0, dev->ev1) == hipSuccess) {
        dev->timed_ms += ms;
        dev->timed_launches++;
    }
    (void)hipGetLastError();
}

extern "C" CFT_API cft_status cftx_kernel_timer(cft_device *dev, int enable)
{
    cft_clear_error();
    if (!dev) {
        cft_set_error("cftx_kernel_timer: dev is NULL");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_status st = dev_enter(dev, "cftx_kernel_timer");
    if (st != CFT_OK)
        return st;
    if (enable && !dev->ev0) {
        if (hipEventCreate(&dev->ev0) != hipSuccess || hipEventCreate(&dev->ev1) != hipSuccess) {
            (void)hipGetLastError();
            cft_set_error("cftx_kernel_timer: creating HIP events failed");
            return CFT_ERR_OUT_OF_MEMORY;
        }
    }
    dev->timing = enable != 0;
    dev->timed_ms = 0.0;
    dev->timed_launches = 0;
    return CFT_OK;
}

extern "C" CFT_API cft_status cftx_kernel_timer_read(cft_device *dev, double *total_ms,
                                                     uint64_t *launches)
{
    cft_clear_error();
    if (!dev || !total_ms || !launches) {
        cft_set_error("cftx_kernel_timer_read: a NULL argument");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *total_ms = dev->timed_ms;
    *launches = dev->timed_launches;
    return CFT_OK;
}

// ---- open / close -----------------------------------------------------------

static cft_status open_impl(const char *artifact, int index, uint32_t depth, cft_device **out,
                            bool own_queue = false)
{
    static const char grammar[] = "the artifact is \"hip:\" followed by the device's number in "
                                  "decimal and nothing else - hip:0 is the first GPU";
    if (!artifact) {
        cft_set_error("cft_open: this library is the MI355X engine and has no software "
                      "backend; open \"hip:N\" (every result it returns is computed on the "
                      "GPU)");
        return CFT_ERR_UNSUPPORTED;
    }
    if (strncmp(artifact, "hip:", 4) != 0) {
        cft_set_error("cft_open: \"%s\" is not a GPU artifact: %s", artifact, grammar);
        return strcmp(artifact, "hip") == 0 ? CFT_ERR_INVALID_ARGUMENT : CFT_ERR_NO_DEVICE;
    }
    const char *p = artifact + 4;
    unsigned long long num = 0;
    bool ok = *p != '\0', big = false;
    for (; ok && *p; p++) {
        if (*p < '0' || *p > '9') {
            ok = false;
            break;
        }
        if (num > (unsigned long long)INT_MAX)
            big = true;
        else
            num = num * 10 + (unsigned)(*p - '0');
    }
    if (!ok) {
        cft_set_error("cft_open: \"%s\" is not a GPU artifact: %s", artifact, grammar);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (index != 0) {
        cft_set_error("cft_open: %s with index %d - hip:N names the device; pass index 0",
                      artifact, index);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        cft_set_error("cft_open: %s: the HIP runtime found no usable GPU%s%s", artifact,
                      e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
        (void)hipGetLastError();
        return CFT_ERR_NO_DEVICE;
    }
    if (big || num >= (unsigned long long)count) {
        cft_set_error("cft_open: %s names a GPU past the %d the HIP runtime sees", artifact,
                      count);
        return CFT_ERR_NO_DEVICE;
    }
#ifdef CFT_PLANTS
    // the negative-control build: CFT_PLANT names the one thing to break
    int plant_fail = 0;
    if (const char *pl = getenv("CFT_PLANT")) {
        if (strcmp(pl, "fail-launch") == 0) {
            plant_fail = 1;
        } else if (*pl) {
            cft_set_error("cft_open: CFT_PLANT=\"%s\" is not a plant of this build (the one "
                          "plant is fail-launch)", pl);
            return CFT_ERR_INVALID_ARGUMENT;
        }
    }
#endif
    cft_device *D = new (std::nothrow) cft_device();
    if (!D)
        return CFT_ERR_OUT_OF_MEMORY;
#ifdef CFT_PLANTS
    D->plant_fail_launch = plant_fail;
#endif
    D->hip_index = (int)num;
    D->scratch_depth = depth;
    hipDeviceProp_t prop;
    if ((e = hipSetDevice(D->hip_index)) != hipSuccess ||
        (e = hipGetDeviceProperties(&prop, D->hip_index)) != hipSuccess) {
        cft_set_error("cft_open: %s: selecting the device failed: %s", artifact,
                      hipGetErrorString(e));
        (void)hipGetLastError();
        delete D;
        return CFT_ERR_NO_DEVICE;
    }
    D->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (prop.sharedMemPerBlock > 0)
        D->lds_max = prop.sharedMemPerBlock;
    // default priority, always: work on streams of different priorities can
    // stall each other indefinitely on this platform.  A stream with a CU
    // mask (here every CU) gets a hardware queue of its own rather than one
    // shared with other streams of the process, so its kernels never wait
    // behind another stream's long launch.
    e = hipErrorUnknown;
    if (own_queue) {
        std::vector<uint32_t> mask((D->num_cus + 31) / 32, 0u);
        for (int c = 0; c < D->num_cus; c++)
            mask[c / 32] |= 1u << (c % 32);
        e = hipExtStreamCreateWithCUMask(&D->stream, (uint32_t)mask.size(), mask.data());
        if (e != hipSuccess)
            (void)hipGetLastError();
    }
    if (e != hipSuccess &&
        (e = hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking)) != hipSuccess) {
        cft_set_error("cft_open: %s: creating a stream failed: %s", artifact,
                      hipGetErrorString(e));
        (void)hipGetLastError();
        delete D;
        return CFT_ERR_NO_DEVICE;
    }
    if (hipMalloc((void **)&D->dwords, 256) != hipSuccess ||
        hipHostMalloc((void **)&D->hwords, 256, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void **)&D->hwords_dev, D->hwords, 0) != hipSuccess ||
        hipMalloc((void **)&D->dslots, CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS * 4) != hipSuccess ||
        hipMemsetAsync(D->dslots, 0, CFTK_FLAG_SLOTS * CFTK_SLOT_WORDS * 4, D->stream) !=
            hipSuccess ||
        hipStreamSynchronize(D->stream) != hipSuccess) {
        (void)hipGetLastError();
        if (D->dwords) (void)hipFree(D->dwords);
        if (D->hwords) (void)hipHostFree(D->hwords);
        if (D->dslots) (void)hipFree(D->dslots);
        (void)hipStreamDestroy(D->stream);
        delete D;
        cft_set_error("cft_open: %s: the handle's flag words could not be allocated",
                      artifact);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    *out = D;
    return CFT_OK;
}

cft_status device_open(int hip_index, uint32_t depth, cft_device **out, bool own_queue)
{
    char name[32];
    snprintf(name, sizeof name, "hip:%d", hip_index);
    return open_impl(name, 0, depth, out, own_queue);
}

extern "C" CFT_API void *cftx_buffer_device_ptr(cft_device *dev, const void *p, size_t bytes)
{
    if (!dev || !p)
        return nullptr;
    cft_buffer *b = find_buffer(dev, p, bytes);
    if (!b || !b->devp)
        return nullptr;
    return (uint8_t *)b->devp + ((const uint8_t *)p - (const uint8_t *)b->host);
}

extern "C" CFT_API cft_status cft_open(const char *artifact, int index, cft_device **out)
{
    cft_clear_error();
    if (!out) {
        cft_set_error("cft_open: out is NULL");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return open_impl(artifact, index, cftl::kDefaultDepth, out);
}

// host/include/cft.h:525-579: a GPU handle opened at a scratch depth, a power
// of two up to 2,048.
extern "C" CFT_API cft_status cft_open_ex(const cft_open_args *args, cft_device **out)
{
    cft_clear_error();
    if (!args || !out) {
        cft_set_error("cft_open_ex: %s is NULL", !args ? "args" : "out");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (args->struct_size != sizeof(cft_open_args)) {
        cft_set_error("cft_open_ex: struct_size is %zu and this library's cft_open_args is "
                      "%zu bytes; an input struct of another size is refused",
                      args->struct_size, sizeof(cft_open_args));
        return CFT_ERR_INVALID_ARGUMENT;
    }
    uint32_t depth = args->scratch_depth ? args->scratch_depth : cftl::kDefaultDepth;
    if ((depth & (depth - 1)) || depth > cftl::kMaxDepth) {
        cft_set_error("cft_open_ex: scratch_depth %u is not a power of two in 1..%u, the "
                      "depths this GPU backend runs",
                      args->scratch_depth, cftl::kMaxDepth);
        return CFT_ERR_INVALID_ARGUMENT;
    }
    return open_impl(args->artifact, args->index, depth, out);
}

extern "C" CFT_API void cft_close(cft_device *dev)
{
    if (!dev)
        return;
    (void)hipSetDevice(dev->hip_index);
    (void)hipStreamSynchronize(dev->stream);
    for (cft_buffer *b : dev->bufs) {
        // a buffer outliving its device stays valid as host memory
        if (b->devp) (void)hipFree(b->devp);
        b->devp = nullptr;
        b->dev = nullptr;
        b->dev_valid = false;
    }
    for (DevArena &a : dev->stage)
        if (a.p) (void)hipFree(a.p);
    for (void *p : dev->retired)
        (void)hipFree(p);
    if (dev->dwords) (void)hipFree(dev->dwords);
    if (dev->hwords) (void)hipHostFree(dev->hwords);
    if (dev->dslots) (void)hipFree(dev->dslots);
    if (dev->ev0) (void)hipEventDestroy(dev->ev0);
    if (dev->ev1) (void)hipEventDestroy(dev->ev1);
    (void)hipStreamDestroy(dev->stream);
    (void)hipGetLastError();
    delete dev;
}

extern "C" CFT_API cft_status cft_get_caps(cft_device *dev, cft_caps *out)
{
    cft_clear_error();
    if (!dev || !out || out->struct_size < sizeof(size_t)) {
        cft_set_error("cft_get_caps: %s", !dev ? "dev is NULL" : !out ? "out is NULL"
                                                : "struct_size is too small to hold itself");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_caps c;
    memset(&c, 0, sizeof c);
    c.format_mask = 0xFu;
    c.tiles = 1;
    c.abi_version = cft_abi_version();
    c.device_version = 0;
    c.flags_readable = 1;
    snprintf(c.backend, sizeof c.backend, "hip-mi355x");
    c.max_deposits = cftl::kMaxDeposits;
    c.max_insns = cftl::kMaxInsns;
    c.max_consts = cftl::kMaxConsts;
    c.seq_features = cftl::kSeqFeatures;
    c.max_scratch = dev->scratch_depth;
    c.buffers_resident = 1;
    size_t want = std::min(out->struct_size, sizeof c);
    c.struct_size = want;
    memcpy(out, &c, want);
    return CFT_OK;
}

// Every assigned opcode in the four formats: the elementwise ones through
// cft_run, the reductions (24, 25, 28, 29, 31) through cft_reduce.
extern "C" CFT_API int cft_supports(cft_device *dev, cft_op op, cft_format fmt)
{
    const int o = (int)op;
    if (!dev || (int)fmt < 0 || (int)fmt > 3)
        return 0;
    const bool assigned = (o >= 0 && o <= 14) || (o >= 16 && o <= 31);
    return assigned ? 1 : 0;
}

// ---- resident buffers (host/include/cft.h:2653-2704, host/src/backend.h:
// 219-279): a pinned mirror the caller writes, a device copy runs bind, and
// the authority rule between them. ----------------------------------------------

extern "C" CFT_API cft_status cft_alloc(cft_device *dev, size_t bytes, cft_buffer **out)
{
    cft_clear_error();
    if (!dev || !out || bytes == 0) {
        cft_set_error("cft_alloc: %s", !dev ? "dev is NULL" : !out ? "out is NULL"
                                             : "a buffer of zero bytes");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    cft_status st = dev_enter(dev, "cft_alloc");
    if (st != CFT_OK)
        return st;
    cft_buffer *b = new (std::nothrow) cft_buffer();
    if (!b)
        return CFT_ERR_OUT_OF_MEMORY;
    if (hipHostMalloc(&b->host, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        cft_set_error("cft_alloc: %zu bytes of pinned host memory could not be allocated",
                      bytes);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    if (hipMalloc(&b->devp, bytes) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(b->host);
        delete b;
        cft_set_error("cft_alloc: %zu bytes of device memory could not be allocated", bytes);
        return CFT_ERR_OUT_OF_MEMORY;
    }
    memset(b->host, 0, bytes);
    b->bytes = bytes;
    b->dev = dev;
    b->host_valid = true;
    b->dev_valid = false;
    dev->bufs.push_back(b);
    *out = b;
    return CFT_OK;
}

extern "C" CFT_API void *cft_buffer_data(cft_buffer *buf) { return buf ? buf->host : nullptr; }

// The mirror is the truth from here.  Refused while the device copy holds a
// run's results that nobody has read back (host/src/backend.h:233-244).
extern "C" CFT_API cft_status cft_buffer_to_device(cft_buffer *buf)
{
    cft_clear_error();
    if (!buf) {
        cft_set_error("cft_buffer_to_device: buf is NULL");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!buf->dev)
        return CFT_OK;                  // its device was closed: host memory
    if (!buf->host_valid) {
        cft_set_error("cft_buffer_to_device: this buffer's device copy holds a run's results "
                      "that have not been read back; call cft_buffer_from_device first");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_status st = dev_enter(buf->dev, "cft_buffer_to_device");
    if (st != CFT_OK)
        return st;
    hipError_t e = hipMemcpyAsync(buf->devp, buf->host, buf->bytes, hipMemcpyHostToDevice,
                                  buf->dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(buf->dev->stream);
    if (e != hipSuccess)
        return dev_fail(buf->dev, e, "cft_buffer_to_device", "the upload");
    buf->dev_valid = true;
    return CFT_OK;
}

extern "C" CFT_API cft_status cft_buffer_from_device(cft_buffer *buf)
{
    cft_clear_error();
    if (!buf) {
        cft_set_error("cft_buffer_from_device: buf is NULL");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    if (!buf->dev || buf->host_valid)
        return CFT_OK;
    cft_status st = dev_enter(buf->dev, "cft_buffer_from_device");
    if (st != CFT_OK)
        return st;
    hipError_t e = hipMemcpyAsync(buf->host, buf->devp, buf->bytes, hipMemcpyDeviceToHost,
                                  buf->dev->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(buf->dev->stream);
    if (e != hipSuccess)
        return dev_fail(buf->dev, e, "cft_buffer_from_device", "the download");
    buf->host_valid = true;
    return CFT_OK;
}

extern "C" CFT_API void cft_buffer_free(cft_buffer *buf)
{
    if (!buf)
        return;
    if (cft_device *dev = buf->dev) {
        (void)hipSetDevice(dev->hip_index);
        (void)hipStreamSynchronize(dev->stream);
        dev->bufs.erase(std::remove(dev->bufs.begin(), dev->bufs.end(), buf), dev->bufs.end());
        if (buf->devp) (void)hipFree(buf->devp);
    }
    (void)hipHostFree(buf->host);
    (void)hipGetLastError();
    delete buf;
}

extern "C" CFT_API cft_status cft_buffer_get_info(cft_buffer *buf, cft_buffer_info *out)
{
    cft_clear_error();
    if (!buf || !out || out->struct_size < sizeof(size_t)) {
        cft_set_error("cft_buffer_get_info: %s", !buf ? "buf is NULL" : !out ? "out is NULL"
                                                  : "struct_size is too small");
        return CFT_ERR_INVALID_ARGUMENT;
    }
    cft_buffer_info info;
    memset(&info, 0, sizeof info);
    info.bytes = buf->bytes;
    info.resident = buf->dev && buf->dev_valid ? 1 : 0;
    info.device_authority = buf->host_valid ? 0 : 1;
    info.resident_binds = buf->resident_binds;
    info.staged_binds = buf->staged_binds;
    snprintf(info.staged_why, sizeof info.staged_why, "%s", buf->why);
    const size_t want = std::min(out->struct_size, sizeof info);
    info.struct_size = want;
    memcpy(out, &info, want);
    return CFT_OK;
}
