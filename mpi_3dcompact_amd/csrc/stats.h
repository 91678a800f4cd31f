// Decaying-HIT statistics of the current state (the turbdata.out hook of
// tests/hit/Samtaney_D4/MPI_3DCompact.cpp:380-556), as device reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "pointwise.h"

namespace mc3d {

constexpr int kStatSums = 11;  // rho mu turbdiss uprime2 uiprime2 uvar sos dUdx^2 rho|u|^2 dil^2 |omega|^2

struct StatsArgs {
    const double *dU[3], *dV[3], *dW[3];  // computational derivatives of U, V, W
    Metrics m;
    const double *rho, *U, *V, *W, *mu, *sos;
    double *partial;   // [blocks][kStatSums]
    double *partial2;  // [blocks]: sum (rho - mean_rho)^2 of the second pass
};
int stats_blocks(long n);
// per-workgroup sums of the pointwise quantities (first pass)
hipError_t launch_hit_stats(const StatsArgs &a, long n, hipStream_t st);
// per-workgroup sums of (rho - mean_rho)^2 (second pass)
hipError_t launch_hit_stats_rho(const double *rho, double mean_rho, double *partial2, long n, hipStream_t st);

}  // namespace mc3d
