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

// gradU[i][j] = d u_i / d x_j from d u_i / d xi_k, src/CurvilinearCSolver_Core.cpp:548-561;
// without metric terms (m.J11 null) the inputs are the Cartesian gradient itself
__device__ __forceinline__ void cart(const Metrics &m, long ip, double d0, double d1, double d2,
                                     double &gx, double &gy, double &gz) {
    if (!m.J11) {
        gx = d0; gy = d1; gz = d2;
        return;
    }
    gx = ldnt(m.J11, ip) * d0 + ldnt(m.J21, ip) * d1 + ldnt(m.J31, ip) * d2;
    gy = ldnt(m.J12, ip) * d0 + ldnt(m.J22, ip) * d1 + ldnt(m.J32, ip) * d2;
    gz = ldnt(m.J13, ip) * d0 + ldnt(m.J23, ip) * d1 + ldnt(m.J33, ip) * d2;
}

// src/DSMSGS.hpp:127-168
__global__ __launch_bounds__(kBlock) void k_dsm_strain(DsmStrainArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    double g00, g01, g02, g10, g11, g12, g20, g21, g22;
    cart(a.m, ip, ldnt(a.dU[0], ip), ldnt(a.dU[1], ip), ldnt(a.dU[2], ip), g00, g01, g02);
    cart(a.m, ip, ldnt(a.dV[0], ip), ldnt(a.dV[1], ip), ldnt(a.dV[2], ip), g10, g11, g12);
    cart(a.m, ip, ldnt(a.dW[0], ip), ldnt(a.dW[1], ip), ldnt(a.dW[2], ip), g20, g21, g22);
    const double S00 = g00, S11 = g11, S22 = g22;
    const double S01 = 0.5 * (g01 + g10);
    const double S02 = 0.5 * (g02 + g20);
    const double S12 = 0.5 * (g12 + g21);
    const double Smag = sqrt(2.0 * (S00 * S00 + S11 * S11 + S22 * S22) +
                             4.0 * (S01 * S01 + S02 * S02 + S12 * S12));
    stnt(a.S[0], ip, S00); stnt(a.S[1], ip, S01); stnt(a.S[2], ip, S02);
    stnt(a.S[3], ip, S11); stnt(a.S[4], ip, S12); stnt(a.S[5], ip, S22);
    stnt(a.Smag, ip, Smag);
    const double rho = ldnt(a.rho, ip);
    stnt(a.M2[0], ip, 2.0 * rho * Smag * (S00 - (1.0 / 3.0) * (S00 + S11 + S22)));
    stnt(a.M2[3], ip, 2.0 * rho * Smag * (S11 - (1.0 / 3.0) * (S00 + S11 + S22)));
    stnt(a.M2[5], ip, 2.0 * rho * Smag * (S22 - (1.0 / 3.0) * (S00 + S11 + S22)));
    stnt(a.M2[1], ip, 2.0 * rho * Smag * S01);
    stnt(a.M2[2], ip, 2.0 * rho * Smag * S02);
    stnt(a.M2[4], ip, 2.0 * rho * Smag * S12);
}

// src/DSMSGS.hpp:198-238
__global__ __launch_bounds__(kBlock) void k_dsm_ml(DsmMLArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rh = ldnt(a.rho_hat, ip), Sh = ldnt(a.Smag_hat, ip);
    const double s00 = ldnt(a.S_hat[0], ip), s01 = ldnt(a.S_hat[1], ip), s02 = ldnt(a.S_hat[2], ip);
    const double s11 = ldnt(a.S_hat[3], ip), s12 = ldnt(a.S_hat[4], ip), s22 = ldnt(a.S_hat[5], ip);
    stnt(a.M[0], ip, ldnt(a.M[0], ip) - 2.0 * kTf2 * rh * Sh * (s00 - (1.0 / 3.0) * (s00 + s11 + s22)));
    stnt(a.M[3], ip, ldnt(a.M[3], ip) - 2.0 * kTf2 * rh * Sh * (s11 - (1.0 / 3.0) * (s00 + s11 + s22)));
    stnt(a.M[5], ip, ldnt(a.M[5], ip) - 2.0 * kTf2 * rh * Sh * (s22 - (1.0 / 3.0) * (s00 + s11 + s22)));
    stnt(a.M[1], ip, ldnt(a.M[1], ip) - 2.0 * kTf2 * rh * Sh * s01);
    stnt(a.M[2], ip, ldnt(a.M[2], ip) - 2.0 * kTf2 * rh * Sh * s02);
    stnt(a.M[4], ip, ldnt(a.M[4], ip) - 2.0 * kTf2 * rh * Sh * s12);
    const double rho = ldnt(a.rho, ip), ru = ldnt(a.rhoU, ip), rv = ldnt(a.rhoV, ip), rw = ldnt(a.rhoW, ip);
    stnt(a.Lt[0], ip, ru * ru / rho);
    stnt(a.Lt[1], ip, ru * rv / rho);
    stnt(a.Lt[2], ip, ru * rw / rho);
    stnt(a.Lt[3], ip, rv * rv / rho);
    stnt(a.Lt[4], ip, rv * rw / rho);
    stnt(a.Lt[5], ip, rw * rw / rho);
}

// src/DSMSGS.hpp:250-270
__global__ __launch_bounds__(kBlock) void k_dsm_l(DsmLArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double uh = ldnt(a.rhoU_hat, ip), vh = ldnt(a.rhoV_hat, ip), wh = ldnt(a.rhoW_hat, ip), rh = ldnt(a.rho_hat, ip);
    stnt(a.L[0], ip, ldnt(a.L[0], ip) - uh * uh / rh);
    stnt(a.L[1], ip, ldnt(a.L[1], ip) - uh * vh / rh);
    stnt(a.L[2], ip, ldnt(a.L[2], ip) - uh * wh / rh);
    stnt(a.L[3], ip, ldnt(a.L[3], ip) - vh * vh / rh);
    stnt(a.L[4], ip, ldnt(a.L[4], ip) - vh * wh / rh);
    stnt(a.L[5], ip, ldnt(a.L[5], ip) - wh * wh / rh);
    if (a.CIdenom) {
        const double Smag = ldnt(a.Smag, ip);
        stnt(a.CIdenom, ip, 2.0 * ldnt(a.rho, ip) * Smag * Smag);
    }
}

// src/DSMSGS.hpp:275-300
__global__ __launch_bounds__(kBlock) void k_dsm_contract(DsmContractArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    if (a.alpha_hat) {
        const double Sh = ldnt(a.Smag_hat, ip);
        stnt(a.CIdenom, ip, 2.0 * kTf2 * ldnt(a.rho_hat, ip) * Sh * Sh - ldnt(a.alpha_hat, ip));
    }
    const double L00 = ldnt(a.L[0], ip), L01 = ldnt(a.L[1], ip), L02 = ldnt(a.L[2], ip);
    const double L11 = ldnt(a.L[3], ip), L12 = ldnt(a.L[4], ip), L22 = ldnt(a.L[5], ip);
    const double M00 = ldnt(a.M[0], ip), M01 = ldnt(a.M[1], ip), M02 = ldnt(a.M[2], ip);
    const double M11 = ldnt(a.M[3], ip), M12 = ldnt(a.M[4], ip), M22 = ldnt(a.M[5], ip);
    const double Lkk = L00 + L11 + L22;
    stnt(a.Lkk, ip, Lkk);
    stnt(a.LijMij, ip, (L00 - (1.0 / 3.0) * Lkk) * M00 + (L11 - (1.0 / 3.0) * Lkk) * M11 +
                           (L22 - (1.0 / 3.0) * Lkk) * M22 + 2.0 * L01 * M01 + 2.0 * L02 * M02 +
                           2.0 * L12 * M12);
    stnt(a.MijMij, ip, M00 * M00 + M11 * M11 + M22 * M22 + 2.0 * M01 * M01 + 2.0 * M02 * M02 +
                           2.0 * M12 * M12);
}

// src/DSMSGS.hpp:314-341
__global__ __launch_bounds__(kBlock) void k_dsm_mu(DsmMuArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rho = ldnt(a.rho, ip), Smag = ldnt(a.Smag, ip);
    stnt(a.mu_sgs, ip, rho * Smag * ldnt(a.LijMij_avg, ip) / ldnt(a.MijMij_avg, ip));
    stnt(a.taukk, ip, a.use_taukk ? 2.0 * rho * Smag * Smag * ldnt(a.Lkk_avg, ip) / ldnt(a.CIdenom_avg, ip) : 0.0);
    if (a.C) {
        const double vol = a.dx * a.dy * a.dz / ldnt(a.J, ip);
        const double del = pow(vol, 1.0 / 3.0);
        stnt(a.C, ip, ldnt(a.LijMij_avg, ip) / (del * del * ldnt(a.MijMij_avg, ip)));
        stnt(a.CI, ip, a.use_taukk ? ldnt(a.Lkk_avg, ip) / (del * del * ldnt(a.CIdenom_avg, ip)) : 0.0);
    }
}

// src/CurvilinearCSolver_Core.cpp:571-584 with the IdealGas relations of src/IdealGas.hpp:50-80
__global__ __launch_bounds__(kBlock) void k_dsm_taukk_eos(const double *rho, const double *taukk,
                                                          Prim pr, Gas g, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double r = ldnt(rho, ip);
    const double p = ldnt(pr.p, ip) - 0.5 * ldnt(taukk, ip);
    const double T = p / (r * g.R_gas);
    stnt(pr.p, ip, p);
    stnt(pr.T, ip, T);
    stnt(pr.mu, ip, member_mu_ref(g, ip) * pow(T / g.T_ref, 0.76));
    stnt(pr.sos, ip, sqrt(g.gamma * p / r));
}

// src/DSMSGS.hpp:345-385
__global__ __launch_bounds__(kBlock) void k_dsm_nk(DsmNKArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    double gx, gy, gz;
    cart(a.m, ip, ldnt(a.dT[0], ip), ldnt(a.dT[1], ip), ldnt(a.dT[2], ip), gx, gy, gz);
    stnt(a.gT[0], ip, gx); stnt(a.gT[1], ip, gy); stnt(a.gT[2], ip, gz);
    const double rho = ldnt(a.rho, ip), Smag = ldnt(a.Smag, ip), T = ldnt(a.T, ip);
    stnt(a.N[0], ip, rho * Smag * gx);
    stnt(a.N[1], ip, rho * Smag * gy);
    stnt(a.N[2], ip, rho * Smag * gz);
    stnt(a.K[0], ip, ldnt(a.rhoU, ip) * T);
    stnt(a.K[1], ip, ldnt(a.rhoV, ip) * T);
    stnt(a.K[2], ip, ldnt(a.rhoW, ip) * T);
    stnt(a.rhoT, ip, rho * T);
}

// src/DSMSGS.hpp:401-425; K_i divides by rho, not rho_hat, as the reference does
__global__ __launch_bounds__(kBlock) void k_dsm_nn(DsmNNArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rh = ldnt(a.rho_hat, ip), Sh = ldnt(a.Smag_hat, ip), rTh = ldnt(a.rhoT_hat, ip), rho = ldnt(a.rho, ip);
    const double N0 = kTf2 * rh * Sh * ldnt(a.g_hat[0], ip) - ldnt(a.NRHS[0], ip);
    const double N1 = kTf2 * rh * Sh * ldnt(a.g_hat[1], ip) - ldnt(a.NRHS[1], ip);
    const double N2 = kTf2 * rh * Sh * ldnt(a.g_hat[2], ip) - ldnt(a.NRHS[2], ip);
    const double K0 = -ldnt(a.rhoU_hat, ip) * rTh / rho + ldnt(a.KRHS[0], ip);
    const double K1 = -ldnt(a.rhoV_hat, ip) * rTh / rho + ldnt(a.KRHS[1], ip);
    const double K2 = -ldnt(a.rhoW_hat, ip) * rTh / rho + ldnt(a.KRHS[2], ip);
    stnt(a.NiNi, ip, N0 * N0 + N1 * N1 + N2 * N2);
    stnt(a.NiKi, ip, K0 * N0 + K1 * N1 + K2 * N2);
}

// src/DSMSGS.hpp:427-444
__global__ __launch_bounds__(kBlock) void k_dsm_ksgs(DsmKArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double NiNi = ldnt(a.NiNi_avg, ip), NiKi = ldnt(a.NiKi_avg, ip);
    const double C_over_Prt = NiKi / NiNi;
    stnt(a.k_sgs, ip, -a.cp * C_over_Prt * ldnt(a.rho, ip) * ldnt(a.Smag, ip));
    if (a.Prt) {
        const double vol = a.dx * a.dy * a.dz / ldnt(a.J, ip);
        const double del = pow(vol, 1.0 / 3.0);
        stnt(a.Prt, ip, -ldnt(a.C, ip) * del * del * NiNi / NiKi);
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
