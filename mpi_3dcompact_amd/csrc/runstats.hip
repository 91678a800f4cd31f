// Running statistics kernels (Stats::updateStatsFields, src/Stats.cpp:121-246).
// Compiled with -ffp-contract=off: every expression is rounded as the
// reference evaluates it.
#include <hip/hip_runtime.h>

#include "pointwise.h"
#include "runstats.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// second moments first (with the previous means), then the means
__global__ __launch_bounds__(kBlock) void k_stats_velocity(StatsArgs s, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double U = ldnt(s.a, ip), V = ldnt(s.b, ip), W = ldnt(s.c, ip);
    const double UA = ldnt(s.f[0], ip), VA = ldnt(s.f[1], ip), WA = ldnt(s.f[2], ip);
    const double UR = ldnt(s.f[3], ip), VR = ldnt(s.f[4], ip), WR = ldnt(s.f[5], ip);
    const double UV = ldnt(s.f[6], ip), UW = ldnt(s.f[7], ip), VW = ldnt(s.f[8], ip);
    const double oW = s.oldW, dt = s.dt, tW = s.totW;
    stnt(s.f[3], ip, (oW * (UR * UR + UA * UA) + dt * U * U) / tW);
    stnt(s.f[4], ip, (oW * (VR * VR + VA * VA) + dt * V * V) / tW);
    stnt(s.f[5], ip, (oW * (WR * WR + WA * WA) + dt * W * W) / tW);
    stnt(s.f[6], ip, (oW * (UV + UA * VA) + dt * U * V) / tW);
    stnt(s.f[7], ip, (oW * (UW + UA * WA) + dt * U * W) / tW);
    stnt(s.f[8], ip, (oW * (VW + VA * WA) + dt * V * W) / tW);
    stnt(s.f[0], ip, (oW * UA + dt * U) / tW);
    stnt(s.f[1], ip, (oW * VA + dt * V) / tW);
    stnt(s.f[2], ip, (oW * WA + dt * W) / tW);
}

__global__ __launch_bounds__(kBlock) void k_stats_velocity_finish(StatsArgs s, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double UA = ldnt(s.f[0], ip), VA = ldnt(s.f[1], ip), WA = ldnt(s.f[2], ip);
    stnt(s.f[3], ip, sqrt(fabs(ldnt(s.f[3], ip) - UA * UA)));
    stnt(s.f[4], ip, sqrt(fabs(ldnt(s.f[4], ip) - VA * VA)));
    stnt(s.f[5], ip, sqrt(fabs(ldnt(s.f[5], ip) - WA * WA)));
    stnt(s.f[6], ip, ldnt(s.f[6], ip) - UA * VA);
    stnt(s.f[7], ip, ldnt(s.f[7], ip) - UA * WA);
    stnt(s.f[8], ip, ldnt(s.f[8], ip) - VA * WA);
}

__global__ __launch_bounds__(kBlock) void k_stats_thermo(StatsArgs s, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double r = ldnt(s.a, ip), p = ldnt(s.b, ip), T = ldnt(s.c, ip);
    const double RA = ldnt(s.f[0], ip), RR = ldnt(s.f[1], ip), PA = ldnt(s.f[2], ip);
    const double PR = ldnt(s.f[3], ip), TA = ldnt(s.f[4], ip), TR = ldnt(s.f[5], ip);
    const double oW = s.oldW, dt = s.dt, tW = s.totW;
    stnt(s.f[1], ip, (oW * (RR * RR + RA * RA) + dt * r * r) / tW);
    stnt(s.f[3], ip, (oW * (PR * PR + PA * PA) + dt * p * p) / tW);
    stnt(s.f[5], ip, (oW * (TR * TR + TA * TA) + dt * T * T) / tW);
    stnt(s.f[0], ip, (oW * RA + dt * r) / tW);
    stnt(s.f[2], ip, (oW * PA + dt * p) / tW);
    stnt(s.f[4], ip, (oW * TA + dt * T) / tW);
}

__global__ __launch_bounds__(kBlock) void k_stats_thermo_finish(StatsArgs s, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double RA = ldnt(s.f[0], ip), PA = ldnt(s.f[2], ip), TA = ldnt(s.f[4], ip);
    stnt(s.f[1], ip, sqrt(fabs(ldnt(s.f[1], ip) - RA * RA)));
    stnt(s.f[3], ip, sqrt(fabs(ldnt(s.f[3], ip) - PA * PA)));
    stnt(s.f[5], ip, sqrt(fabs(ldnt(s.f[5], ip) - TA * TA)));
}

}  // namespace

hipError_t launch_stats_velocity(const StatsArgs &s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_velocity, grid_for(n), dim3(kBlock), 0, st, s, n);
    return hipGetLastError();
}
hipError_t launch_stats_velocity_finish(const StatsArgs &s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_velocity_finish, grid_for(n), dim3(kBlock), 0, st, s, n);
    return hipGetLastError();
}
hipError_t launch_stats_thermo(const StatsArgs &s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_thermo, grid_for(n), dim3(kBlock), 0, st, s, n);
    return hipGetLastError();
}
hipError_t launch_stats_thermo_finish(const StatsArgs &s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_thermo_finish, grid_for(n), dim3(kBlock), 0, st, s, n);
    return hipGetLastError();
}

}  // namespace mc3d
