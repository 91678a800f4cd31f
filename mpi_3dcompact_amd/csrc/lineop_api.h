// Internal entry of the line operators used by the host classes.
#pragma once

#include "mc3d.h"

namespace mc3d {

// Like mc3d_deriv_apply / mc3d_filter_apply, optionally adding the result
// into out instead of overwriting it.
int lineop_launch_accumulate(const mc3d_lineop *op, const double *const *in, double *const *out,
                             int nfields, int accumulate, void *stream);

// Symbol-style name of the kernel instantiation an operator launches,
// e.g. "lineop_kernel<2, 1, 0>" (half bandwidth, direction, boundary mode).
const char *lineop_kernel_name(const mc3d_lineop *op);

}  // namespace mc3d
