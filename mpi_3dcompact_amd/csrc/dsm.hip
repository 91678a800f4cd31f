// Pointwise stages of the dynamic Smagorinsky model (src/DSMSGS.hpp:124-444).
//
// HBM-bound streaming kernels, one thread per point, compiled with
// -ffp-contract=off: every expression keeps the reference's evaluation order
// with separately rounded products and sums.  The Cartesian gradients are
// formed on the fly from the computational derivatives and the metrics,
// exactly as the flux kernel forms them.
#include <hip/hip_runtime.h>

#include "dsm.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;
constexpr double kTf2 = 2.0 * 2.0;  // test-filter width ratio squared, src/DSMSGS.hpp:55

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// gradU[i][j] = d u_i / d x_j from d u_i / d xi_k, src/CurvilinearCSolver_Core.cpp:548-561
__device__ __forceinline__ void cart(const Metrics &m, long ip, double d0, double d1, double d2,
                                     double &gx, double &gy, double &gz) {
    gx = m.J11[ip] * d0 + m.J21[ip] * d1 + m.J31[ip] * d2;
    gy = m.J12[ip] * d0 + m.J22[ip] * d1 + m.J32[ip] * d2;
    gz = m.J13[ip] * d0 + m.J23[ip] * d1 + m.J33[ip] * d2;
}

// src/DSMSGS.hpp:127-168
__global__ __launch_bounds__(kBlock) void k_dsm_strain(DsmStrainArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    double g00, g01, g02, g10, g11, g12, g20, g21, g22;
    cart(a.m, ip, a.dU[0][ip], a.dU[1][ip], a.dU[2][ip], g00, g01, g02);
    cart(a.m, ip, a.dV[0][ip], a.dV[1][ip], a.dV[2][ip], g10, g11, g12);
    cart(a.m, ip, a.dW[0][ip], a.dW[1][ip], a.dW[2][ip], g20, g21, g22);
    const double S00 = g00, S11 = g11, S22 = g22;
    const double S01 = 0.5 * (g01 + g10);
    const double S02 = 0.5 * (g02 + g20);
    const double S12 = 0.5 * (g12 + g21);
    const double Smag = sqrt(2.0 * (S00 * S00 + S11 * S11 + S22 * S22) +
                             4.0 * (S01 * S01 + S02 * S02 + S12 * S12));
    a.S[0][ip] = S00; a.S[1][ip] = S01; a.S[2][ip] = S02;
    a.S[3][ip] = S11; a.S[4][ip] = S12; a.S[5][ip] = S22;
    a.Smag[ip] = Smag;
    const double rho = a.rho[ip];
    a.M2[0][ip] = 2.0 * rho * Smag * (S00 - (1.0 / 3.0) * (S00 + S11 + S22));
    a.M2[3][ip] = 2.0 * rho * Smag * (S11 - (1.0 / 3.0) * (S00 + S11 + S22));
    a.M2[5][ip] = 2.0 * rho * Smag * (S22 - (1.0 / 3.0) * (S00 + S11 + S22));
    a.M2[1][ip] = 2.0 * rho * Smag * S01;
    a.M2[2][ip] = 2.0 * rho * Smag * S02;
    a.M2[4][ip] = 2.0 * rho * Smag * S12;
}

// src/DSMSGS.hpp:198-238
__global__ __launch_bounds__(kBlock) void k_dsm_ml(DsmMLArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rh = a.rho_hat[ip], Sh = a.Smag_hat[ip];
    const double s00 = a.S_hat[0][ip], s01 = a.S_hat[1][ip], s02 = a.S_hat[2][ip];
    const double s11 = a.S_hat[3][ip], s12 = a.S_hat[4][ip], s22 = a.S_hat[5][ip];
    a.M[0][ip] = a.M[0][ip] - 2.0 * kTf2 * rh * Sh * (s00 - (1.0 / 3.0) * (s00 + s11 + s22));
    a.M[3][ip] = a.M[3][ip] - 2.0 * kTf2 * rh * Sh * (s11 - (1.0 / 3.0) * (s00 + s11 + s22));
    a.M[5][ip] = a.M[5][ip] - 2.0 * kTf2 * rh * Sh * (s22 - (1.0 / 3.0) * (s00 + s11 + s22));
    a.M[1][ip] = a.M[1][ip] - 2.0 * kTf2 * rh * Sh * s01;
    a.M[2][ip] = a.M[2][ip] - 2.0 * kTf2 * rh * Sh * s02;
    a.M[4][ip] = a.M[4][ip] - 2.0 * kTf2 * rh * Sh * s12;
    const double rho = a.rho[ip], ru = a.rhoU[ip], rv = a.rhoV[ip], rw = a.rhoW[ip];
    a.Lt[0][ip] = ru * ru / rho;
    a.Lt[1][ip] = ru * rv / rho;
    a.Lt[2][ip] = ru * rw / rho;
    a.Lt[3][ip] = rv * rv / rho;
    a.Lt[4][ip] = rv * rw / rho;
    a.Lt[5][ip] = rw * rw / rho;
}

// src/DSMSGS.hpp:250-270
__global__ __launch_bounds__(kBlock) void k_dsm_l(DsmLArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double uh = a.rhoU_hat[ip], vh = a.rhoV_hat[ip], wh = a.rhoW_hat[ip], rh = a.rho_hat[ip];
    a.L[0][ip] = a.L[0][ip] - uh * uh / rh;
    a.L[1][ip] = a.L[1][ip] - uh * vh / rh;
    a.L[2][ip] = a.L[2][ip] - uh * wh / rh;
    a.L[3][ip] = a.L[3][ip] - vh * vh / rh;
    a.L[4][ip] = a.L[4][ip] - vh * wh / rh;
    a.L[5][ip] = a.L[5][ip] - wh * wh / rh;
    if (a.CIdenom) {
        const double Smag = a.Smag[ip];
        a.CIdenom[ip] = 2.0 * a.rho[ip] * Smag * Smag;
    }
}

// src/DSMSGS.hpp:275-300
__global__ __launch_bounds__(kBlock) void k_dsm_contract(DsmContractArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    if (a.alpha_hat) {
        const double Sh = a.Smag_hat[ip];
        a.CIdenom[ip] = 2.0 * kTf2 * a.rho_hat[ip] * Sh * Sh - a.alpha_hat[ip];
    }
    const double L00 = a.L[0][ip], L01 = a.L[1][ip], L02 = a.L[2][ip];
    const double L11 = a.L[3][ip], L12 = a.L[4][ip], L22 = a.L[5][ip];
    const double M00 = a.M[0][ip], M01 = a.M[1][ip], M02 = a.M[2][ip];
    const double M11 = a.M[3][ip], M12 = a.M[4][ip], M22 = a.M[5][ip];
    const double Lkk = L00 + L11 + L22;
    a.Lkk[ip] = Lkk;
    a.LijMij[ip] = (L00 - (1.0 / 3.0) * Lkk) * M00 + (L11 - (1.0 / 3.0) * Lkk) * M11 +
                   (L22 - (1.0 / 3.0) * Lkk) * M22 + 2.0 * L01 * M01 + 2.0 * L02 * M02 +
                   2.0 * L12 * M12;
    a.MijMij[ip] = M00 * M00 + M11 * M11 + M22 * M22 + 2.0 * M01 * M01 + 2.0 * M02 * M02 +
                   2.0 * M12 * M12;
}

// src/DSMSGS.hpp:314-341
__global__ __launch_bounds__(kBlock) void k_dsm_mu(DsmMuArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rho = a.rho[ip], Smag = a.Smag[ip];
    a.mu_sgs[ip] = rho * Smag * a.LijMij_avg[ip] / a.MijMij_avg[ip];
    a.taukk[ip] = a.use_taukk ? 2.0 * rho * Smag * Smag * a.Lkk_avg[ip] / a.CIdenom_avg[ip] : 0.0;
    if (a.C) {
        const double vol = a.dx * a.dy * a.dz / a.J[ip];
        const double del = pow(vol, 1.0 / 3.0);
        a.C[ip] = a.LijMij_avg[ip] / (del * del * a.MijMij_avg[ip]);
        a.CI[ip] = a.use_taukk ? a.Lkk_avg[ip] / (del * del * a.CIdenom_avg[ip]) : 0.0;
    }
}

// src/CurvilinearCSolver_Core.cpp:571-584 with the IdealGas relations of src/IdealGas.hpp:50-80
__global__ __launch_bounds__(kBlock) void k_dsm_taukk_eos(const double *rho, const double *taukk,
                                                          Prim pr, Gas g, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double r = rho[ip];
    const double p = pr.p[ip] - 0.5 * taukk[ip];
    const double T = p / (r * g.R_gas);
    pr.p[ip] = p;
    pr.T[ip] = T;
    pr.mu[ip] = member_mu_ref(g, ip) * pow(T / g.T_ref, 0.76);
    pr.sos[ip] = sqrt(g.gamma * p / r);
}

// src/DSMSGS.hpp:345-385
__global__ __launch_bounds__(kBlock) void k_dsm_nk(DsmNKArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    double gx, gy, gz;
    cart(a.m, ip, a.dT[0][ip], a.dT[1][ip], a.dT[2][ip], gx, gy, gz);
    a.gT[0][ip] = gx; a.gT[1][ip] = gy; a.gT[2][ip] = gz;
    const double rho = a.rho[ip], Smag = a.Smag[ip], T = a.T[ip];
    a.N[0][ip] = rho * Smag * gx;
    a.N[1][ip] = rho * Smag * gy;
    a.N[2][ip] = rho * Smag * gz;
    a.K[0][ip] = a.rhoU[ip] * T;
    a.K[1][ip] = a.rhoV[ip] * T;
    a.K[2][ip] = a.rhoW[ip] * T;
    a.rhoT[ip] = rho * T;
}

// src/DSMSGS.hpp:401-425; K_i divides by rho, not rho_hat, as the reference does
__global__ __launch_bounds__(kBlock) void k_dsm_nn(DsmNNArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rh = a.rho_hat[ip], Sh = a.Smag_hat[ip], rTh = a.rhoT_hat[ip], rho = a.rho[ip];
    const double N0 = kTf2 * rh * Sh * a.g_hat[0][ip] - a.NRHS[0][ip];
    const double N1 = kTf2 * rh * Sh * a.g_hat[1][ip] - a.NRHS[1][ip];
    const double N2 = kTf2 * rh * Sh * a.g_hat[2][ip] - a.NRHS[2][ip];
    const double K0 = -a.rhoU_hat[ip] * rTh / rho + a.KRHS[0][ip];
    const double K1 = -a.rhoV_hat[ip] * rTh / rho + a.KRHS[1][ip];
    const double K2 = -a.rhoW_hat[ip] * rTh / rho + a.KRHS[2][ip];
    a.NiNi[ip] = N0 * N0 + N1 * N1 + N2 * N2;
    a.NiKi[ip] = K0 * N0 + K1 * N1 + K2 * N2;
}

// src/DSMSGS.hpp:427-444
__global__ __launch_bounds__(kBlock) void k_dsm_ksgs(DsmKArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double NiNi = a.NiNi_avg[ip], NiKi = a.NiKi_avg[ip];
    const double C_over_Prt = NiKi / NiNi;
    a.k_sgs[ip] = -a.cp * C_over_Prt * a.rho[ip] * a.Smag[ip];
    if (a.Prt) {
        const double vol = a.dx * a.dy * a.dz / a.J[ip];
        const double del = pow(vol, 1.0 / 3.0);
        a.Prt[ip] = -a.C[ip] * del * del * NiNi / NiKi;
    }
}

}  // namespace

hipError_t launch_dsm_strain(const DsmStrainArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_strain, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_ml(const DsmMLArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_ml, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_l(const DsmLArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_l, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_contract(const DsmContractArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_contract, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_mu(const DsmMuArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_mu, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_taukk_eos(const double *rho, const double *taukk, Prim pr, Gas g, long n,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_taukk_eos, grid_for(n), dim3(kBlock), 0, st, rho, taukk, pr, g, n);
    return hipGetLastError();
}
hipError_t launch_dsm_nk(const DsmNKArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_nk, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_nn(const DsmNNArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_nn, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}
hipError_t launch_dsm_ksgs(const DsmKArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dsm_ksgs, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}

}  // namespace mc3d
