// Decaying-HIT statistics (tests/hit/Samtaney_D4/MPI_3DCompact.cpp:380-556):
// pointwise quantities reduced per workgroup in a fixed order (FP64); the
// host adds the workgroup sums in order, so results do not depend on timing.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "stats.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 4096;

__device__ __forceinline__ double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// :413-460: Cartesian gradient, dilatation, vorticity, traceless strain, the
// resolved dissipation 2 mu S*_ij du_i/dx_j and the velocity moments
__global__ __launch_bounds__(kBlock) void k_hit_stats(StatsArgs a, long n) {
    __shared__ double sh[kBlock];
    double acc[kStatSums];
    for (int k = 0; k < kStatSums; ++k) acc[k] = 0.0;
    for (long p = (long)blockIdx.x * kBlock + threadIdx.x; p < n; p += (long)gridDim.x * kBlock) {
        double g[3][3];
        const double *const *d[3] = {a.dU, a.dV, a.dW};
        const double J11 = a.m.J11[p], J12 = a.m.J12[p], J13 = a.m.J13[p];
        const double J21 = a.m.J21[p], J22 = a.m.J22[p], J23 = a.m.J23[p];
        const double J31 = a.m.J31[p], J32 = a.m.J32[p], J33 = a.m.J33[p];
        for (int i = 0; i < 3; ++i) {
            const double d1 = d[i][0][p], d2 = d[i][1][p], d3 = d[i][2][p];
            g[i][0] = J11 * d1 + J21 * d2 + J31 * d3;
            g[i][1] = J12 * d1 + J22 * d2 + J32 * d3;
            g[i][2] = J13 * d1 + J23 * d2 + J33 * d3;
        }
        const double dil = g[0][0] + g[1][1] + g[2][2];
        const double wx = g[2][1] - g[1][2], wy = g[0][2] - g[2][0], wz = g[1][0] - g[0][1];
        double diss = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double s = 0.5 * (g[i][j] + g[j][i]) - (i == j ? (1.0 / 3.0) * dil : 0.0);
                diss = diss + s * g[i][j];
            }
        const double u = a.U[p], v = a.V[p], w = a.W[p], r = a.rho[p];
        const double uvar = u * u + v * v + w * w;
        acc[0] += r;
        acc[1] += a.mu[p];
        acc[2] += 2.0 * a.mu[p] * diss;
        acc[3] += uvar / 3.0;
        acc[4] += (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2]) / 3.0;
        acc[5] += uvar;
        acc[6] += a.sos[p];
        acc[7] += g[0][0] * g[0][0];
        acc[8] += r * uvar;
        acc[9] += dil * dil;
        acc[10] += wx * wx + wy * wy + wz * wz;
    }
    for (int k = 0; k < kStatSums; ++k) {
        const double s = block_sum(acc[k], sh);
        if (threadIdx.x == 0) a.partial[(size_t)blockIdx.x * kStatSums + k] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_hit_stats_rho(const double *rho, double mean, double *partial2, long n) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (long p = (long)blockIdx.x * kBlock + threadIdx.x; p < n; p += (long)gridDim.x * kBlock) {
        const double d = rho[p] - mean;
        acc += d * d;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial2[blockIdx.x] = s;
}

}  // namespace

int stats_blocks(long n) { return (int)std::max<long>(1, std::min<long>((n + kBlock - 1) / kBlock, kMaxBlocks)); }

hipError_t launch_hit_stats(const StatsArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_hit_stats, dim3((unsigned)stats_blocks(n)), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_hit_stats_rho(const double *rho, double mean_rho, double *partial2, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_hit_stats_rho, dim3((unsigned)stats_blocks(n)), dim3(kBlock), 0, st, rho, mean_rho,
                       partial2, n);
    return hipGetLastError();
}

}  // namespace mc3d
