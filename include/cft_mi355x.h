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

#ifdef __cplusplus
}
#endif

#endif /* CFT_MI355X_H */
