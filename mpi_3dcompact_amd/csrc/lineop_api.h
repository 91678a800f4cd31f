// Internal entry of the line operators used by the host classes.
#pragma once

#include "lineop.h"
#include "mc3d.h"

namespace mc3d {

// Like mc3d_deriv_apply / mc3d_filter_apply, optionally adding the result
// into out instead of overwriting it.
int lineop_launch_accumulate(const mc3d_lineop *op, const double *const *in, double *const *out,
                             int nfields, int accumulate, void *stream);

// Launch shape of a line operator: rows per chunk (16 or 32; 0 chooses),
// kernel family (0 chooses: the fast kernels when they take the operator;
// 1 the fast kernels; 2 the LDS-tiled kernel) and lines per workgroup of the
// LDS-tiled kernel (a power of two, 0 chooses).  Every shape gives the same
// operator; the rounding of the partitioned solve depends on the chunking.
struct LineTune {
    int mm = 0;
    int family = 0;
    int lines = 0;
};

// Operator over nbatch members of box dims stacked along z (periodic lines
// in the directions across the operator's when nbatch > 1); cls kDeriv
// (kind MC3D_CD2..PENTA10, spacing dd) or kFilter (kind MC3D_COMPACT8..).
int lineop_create(int cls, int kind, int dir, const int bc3[3], const int dims[3], double dd, double alphaF,
                  int nbatch, const LineTune &tune, mc3d_lineop **out);
// the shape an operator was built with (family 1 or 2, the chunk rows, lines)
LineTune lineop_shape(const mc3d_lineop *op);

// Symbol-style name of the kernel instantiation an operator launches,
// e.g. "lineop_kernel<2, 1, 0>" (half bandwidth, direction, boundary mode).
const char *lineop_kernel_name(const mc3d_lineop *op);

}  // namespace mc3d
