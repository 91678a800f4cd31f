// Wall fix-ups of the solver step (src/CurvilinearCSolver_BC.cpp:63-322 at the
// initial condition, :372-700 at every stage), one launch per face so that
// points on edges see the faces in the reference's order.  Compiled with
// -ffp-contract=off: rounded as the reference rounds.
#include <hip/hip_runtime.h>

#include "bc.h"
#include "mc3d.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double neumann(int scheme, const double f[6]) {
    if (scheme == MC3D_CD2) return (f[0] * 4.0 - f[1] * 1.0) / 3.0;
    return (f[0] * 360.0 - f[1] * 450.0 + f[2] * 400.0 - f[3] * 225.0 + f[4] * 72.0 - f[5] * 10.0) / 147.0;
}

__global__ void k_wall_face(WallArgs a, long npts) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    const int d = a.face / 2, hi = a.face % 2;
    const long sx = 1, sy = a.Nx, sz = (long)a.Nx * a.Ny;
    long ip, step;
    if (d == 0) {
        const long j = t % a.Ny, k = t / a.Ny;
        ip = (hi ? a.Nx - 1 : 0) * sx + j * sy + k * sz;
        step = hi ? -sx : sx;
    } else if (d == 1) {
        const long i = t % a.Nx, k = t / a.Nx;
        ip = i * sx + (hi ? a.Ny - 1 : 0) * sy + k * sz;
        step = hi ? -sy : sy;
    } else {
        const long i = t % a.Nx, j = t / a.Nx;
        ip = i * sx + j * sy + (hi ? a.Nz - 1 : 0) * sz;
        step = hi ? -sz : sz;
    }
    const Prim &pr = a.pr;
    pr.U[ip] = 0.0;
    pr.V[ip] = 0.0;
    pr.W[ip] = 0.0;
    const double rho = a.q.q[0][ip];
    if (a.moving) {
        a.q.q[1][ip] = rho * 0.0;
        a.q.q[2][ip] = rho * 0.0;
        a.q.q[3][ip] = rho * 0.0;
    } else {
        a.q.q[1][ip] = 0.0;
        a.q.q[2][ip] = 0.0;
        a.q.q[3][ip] = 0.0;
    }
    if (a.neumann) {
        double f[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) f[m] = pr.T[ip + (m + 1) * step];
        pr.T[ip] = neumann(a.scheme, f);
    }
    if (a.stage) {
        const double p = rho * a.g.R_gas * pr.T[ip];
        pr.p[ip] = p;
        a.q.q[4][ip] = p / (a.g.gamma - 1.0) + 0.5 * rho * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    }
}

__global__ void k_wall_ic_update(Cons r1, Prim pr, Gas g, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rho = r1.q[0][ip];
    const double p = rho * g.R_gas * pr.T[ip];
    pr.p[ip] = p;
    pr.sos[ip] = sqrt(g.gamma * p / rho);
    const double U = pr.U[ip], V = pr.V[ip], W = pr.W[ip];
    r1.q[4][ip] = p / (g.gamma - 1.0) + 0.5 * rho * (U * U + V * V + W * W);
}

}  // namespace

hipError_t launch_wall_face(const WallArgs &a, hipStream_t st) {
    if (a.face < 0 || a.face > 5) return hipErrorInvalidValue;
    const int d = a.face / 2;
    const long npts = d == 0 ? (long)a.Ny * a.Nz : d == 1 ? (long)a.Nx * a.Nz : (long)a.Nx * a.Ny;
    const int nline = d == 0 ? a.Nx : d == 1 ? a.Ny : a.Nz;
    if (a.neumann && nline < 7) return hipErrorInvalidValue;  // six inward points
    hipLaunchKernelGGL(k_wall_face, dim3((unsigned)((npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a, npts);
    return hipGetLastError();
}

hipError_t launch_wall_ic_update(Cons r1, Prim pr, Gas g, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_wall_ic_update, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, r1, pr,
                       g, n);
    return hipGetLastError();
}

}  // namespace mc3d
