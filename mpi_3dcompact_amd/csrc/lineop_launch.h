// Launch interface of the line-operator kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lineop.h"

namespace mc3d {

enum LaunchMode { kModePlainPeriodic = 0, kModePlainDirichlet = 1, kModeTPB = 2 };

struct LaunchArgs {
    const double *in[kMaxFields];
    double *out[kMaxFields];
    const double *wrap[6];  // TPB wrap values {Nm3, Nm2, Nm1, Np1, Np2, Np3}
    int Nx, Ny, Nz;  // one member's box
    // members stacked along z (>= 1).  x- and y-lines run over all of them as
    // over one box nbatch*Nz tall (nlines counts them all); z-lines are
    // launched with the member as grid z (nlines counts one member's lines)
    int nbatch;
    int nlines;
    int accumulate;  // out += result instead of out = result
};

// LDS-tiled kernel, any chunking, all modes (TPB included)
size_t lineop_lds_bytes(const OpDev &op);
hipError_t lineop_launch(const OpDev &op, const LaunchArgs &a, int nfields, int mode,
                         hipStream_t st);

// fast kernels: evenly chunked lines (N = P * MM), plain periodic (no special
// rows) or Dirichlet with dense special rows; no TPB
size_t lineop_fast_lds_bytes(const OpDev &op);
bool lineop_fast_supported(const OpDev &op, size_t lds_limit);
// groups of chunks the 64-line x-line kernel stages one after another so that
// its tile fits lds_limit (1: the whole line at once)
int lineop_fast_phases(const OpDev &op, size_t lds_limit);
// Lines the fast path runs two chunks to a wave: y- and z-lines
// (lineop_strided_pair) evenly chunked into 17 to 32 chunks whose chunks read
// chunk-independent tables (periodic lines without special rows), or, up to
// 30 chunks, need none or have interior chunks sharing theirs (Dirichlet
// lines, CD2); x-lines (lineop_contig_pair) of banded Dirichlet operators in
// 17 to 30 chunks of 32 rows.  And the waves of its workgroup.
bool lineop_fast_pair(const OpDev &op);
int lineop_pair_waves(const OpDev &op);
hipError_t lineop_launch_fast(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);

}  // namespace mc3d
