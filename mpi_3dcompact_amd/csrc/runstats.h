// Running statistics of the solution (src/Stats.hpp, src/Stats.cpp): device
// kernels of Stats::updateStatsFields.
#pragma once

#include <hip/hip_runtime.h>

namespace mc3d {

// Field order, as the statistics files hold them (src/Stats.cpp:5-22,62-76):
// velocity UAVG VAVG WAVG URMS VRMS WRMS UVREY UWREY VWREY; thermo RHOAVG
// RHORMS PAVG PRMS TAVG TRMS.
struct StatsArgs {
    const double *a, *b, *c;  // U V W, or rho2 p T
    double *f[9];
    double oldW, dt, totW;
};

// time-weighted accumulation, before the line averages (src/Stats.cpp:128-146, 186-198)
hipError_t launch_stats_velocity(const StatsArgs &s, long n, hipStream_t st);
hipError_t launch_stats_thermo(const StatsArgs &s, long n, hipStream_t st);
// RMS and Reynolds stresses from the averaged moments (:158-167, 207-212)
hipError_t launch_stats_velocity_finish(const StatsArgs &s, long n, hipStream_t st);
hipError_t launch_stats_thermo_finish(const StatsArgs &s, long n, hipStream_t st);

}  // namespace mc3d
