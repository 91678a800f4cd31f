// Metric terms from mesh coordinates: the pointwise stages and the wrap
// values of the transformed-periodic derivatives (the derivatives themselves
// are the line operators).  Compiled with -ffp-contract=off: every product
// and sum is rounded as the reference rounds it.  Init-time kernels, one
// thread per point (per line for the wrap values).
#include <hip/hip_runtime.h>

#include "metrics.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// line geometry of direction DIR (line order of the line operators); a
// compile-time direction keeps the three cases apart in the generated code
template <int DIR>
__device__ __forceinline__ void line_geometry(long q, int Nx, int Ny, int Nz, long &base, long &stride, int &n) {
    if (DIR == 0) { base = q * Nx; stride = 1; n = Nx; }
    if (DIR == 1) { base = (q / Nx) * (long)Nx * Ny + q % Nx; stride = Nx; n = Ny; }
    if (DIR == 2) { base = q; stride = (long)Nx * Ny; n = Nz; }
}

template <int DIR>
__global__ void k_tpb_wraps(WrapArgs a, int Nx, int Ny, int Nz, long nlines) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nlines) return;
    long base, stride;
    int n;
    line_geometry<DIR>(q, Nx, Ny, Nz, base, stride, n);
    for (int m = 0; m < 6; ++m) {
        const bool lo = m < 3;  // Nm slots: the line's last three points, shifted down
        const long p = base + (lo ? (n - 3 + m) : (m - 3)) * stride;
        double v = 0.0;
        for (int t = 0; t < a.nterms; ++t) {
            const double s = lo ? a.A[t][p] - a.T[t] : a.A[t][p] + a.T[t];
            const double term = a.B[t] ? s * a.B[t][p] : s;
            v = t == 0 ? term : v + term;
        }
        a.out[m][q] = v;
    }
}

__global__ void k_metric_products(const double *x11, const double *x12, const double *x13,
                                  const double *x21, const double *x22, const double *x23,
                                  const double *y, const double *z, double *Va1, double *Va2,
                                  double *Va3, double *Vb1, double *Vb2, double *Vb3, double *Vc1,
                                  double *Vc2, double *Vc3, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double Y = y[p], Z = z[p];
    Va1[p] = x11[p] * Y; Va2[p] = x12[p] * Y; Va3[p] = x13[p] * Y;
    Vb1[p] = x11[p] * Z; Vb2[p] = x12[p] * Z; Vb3[p] = x13[p] * Z;
    Vc1[p] = x21[p] * Z; Vc2[p] = x22[p] * Z; Vc3[p] = x23[p] * Z;
}

__global__ void k_metric_cofactors(CofactorArgs a, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double J[3][3];
    J[0][0] = a.dVc[1][2][p] - a.dVc[2][1][p];
    J[0][1] = a.dVb[2][1][p] - a.dVb[1][2][p];
    J[0][2] = a.dVa[1][2][p] - a.dVa[2][1][p];
    J[1][0] = a.dVc[2][0][p] - a.dVc[0][2][p];
    J[1][1] = a.dVb[0][2][p] - a.dVb[2][0][p];
    J[1][2] = a.dVa[2][0][p] - a.dVa[0][2][p];
    J[2][0] = a.dVc[0][1][p] - a.dVc[1][0][p];
    J[2][1] = a.dVb[1][0][p] - a.dVb[0][1][p];
    J[2][2] = a.dVa[0][1][p] - a.dVa[1][0][p];
    const double x = a.x[p], y = a.y[p], z = a.z[p];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) a.Jm[i][j][p] = J[i][j];
        a.pre[i][p] = x * J[i][0] + y * J[i][1] + z * J[i][2];
    }
}

struct Ptr9 {
    double *p[9];
};

__global__ void k_metric_final9(const double *J1, const double *J2, const double *J3, double *J,
                                Ptr9 m, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double j = 1 / ((1.0 / 3.0) * (J1[p] + J2[p] + J3[p]));
    J[p] = j;
#pragma unroll
    for (int k = 0; k < 9; ++k) m.p[k][p] = j * m.p[k][p];
}

}  // namespace

hipError_t launch_tpb_wraps(const WrapArgs &a, int dir, int Nx, int Ny, int Nz, hipStream_t st) {
    if (dir < 0 || dir > 2) return hipErrorInvalidValue;
    const int n = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
    const long nlines = (long)Nx * Ny * Nz / n;
    if (dir == 0) hipLaunchKernelGGL(k_tpb_wraps<0>, grid_for(nlines), dim3(kBlock), 0, st, a, Nx, Ny, Nz, nlines);
    if (dir == 1) hipLaunchKernelGGL(k_tpb_wraps<1>, grid_for(nlines), dim3(kBlock), 0, st, a, Nx, Ny, Nz, nlines);
    if (dir == 2) hipLaunchKernelGGL(k_tpb_wraps<2>, grid_for(nlines), dim3(kBlock), 0, st, a, Nx, Ny, Nz, nlines);
    return hipGetLastError();
}

hipError_t launch_metric_products(const double *const xE1[3], const double *const xE2[3], const double *y,
                                  const double *z, double *const Va[3], double *const Vb[3],
                                  double *const Vc[3], long n, hipStream_t st) {
    hipLaunchKernelGGL(k_metric_products, grid_for(n), dim3(kBlock), 0, st, xE1[0], xE1[1], xE1[2], xE2[0],
                       xE2[1], xE2[2], y, z, Va[0], Va[1], Va[2], Vb[0], Vb[1], Vb[2], Vc[0], Vc[1], Vc[2], n);
    return hipGetLastError();
}

hipError_t launch_metric_cofactors(const CofactorArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_metric_cofactors, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_metric_final(const double *const Jdet[3], double *J, double *const Jm[9], long n,
                               hipStream_t st) {
    Ptr9 m;
    for (int k = 0; k < 9; ++k) m.p[k] = Jm[k];
    hipLaunchKernelGGL(k_metric_final9, grid_for(n), dim3(kBlock), 0, st, Jdet[0], Jdet[1], Jdet[2], J, m, n);
    return hipGetLastError();
}

}  // namespace mc3d
