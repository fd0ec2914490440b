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

#ifdef __cplusplus
}
#endif

#endif /* CFT_MI355X_H */
