// Pointwise kernels of the solver step: initial conversion, equation of state,
// CFL reduction, curvilinear flux assembly, RHS + Runge-Kutta register update,
// Vreman eddy viscosity.
//
// All are HBM-bound streaming kernels.  This file is compiled with
// -ffp-contract=off: separately rounded products and sums, in the reference's
// evaluation order, make every pointwise result reproduce the reference's
// rounding; FMA buys nothing at this arithmetic intensity.
#include <hip/hip_runtime.h>

#include "pointwise.h"

namespace mc3d {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

__device__ __forceinline__ double solve_mu(const Gas &g, double T, long ip) {
    return member_mu_ref(g, ip) * pow(T / g.T_ref, 0.76);
}
__device__ __forceinline__ double solve_p(const Gas &g, double rho, double rhoE, double U, double V,
                                          double W) {
    return (g.gamma - 1) * (rhoE - 0.5 * rho * (U * U + V * V + W * W));
}
__device__ __forceinline__ double solve_T(const Gas &g, double rho, double p) {
    return p / (rho * g.R_gas);
}
__device__ __forceinline__ double solve_sos(const Gas &g, double rho, double p) {
    return sqrt(g.gamma * p / rho);
}

// setInitialConditions, src/CurvilinearCSolver_BC.cpp:30-63
__global__ void k_init_state(const double *__restrict__ rho0, const double *__restrict__ U0,
                             const double *__restrict__ V0, const double *__restrict__ W0,
                             const double *__restrict__ p0, Cons r1, Prim pr, Metrics m, Gas g,
                             long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double U = U0[ip], V = V0[ip], W = W0[ip], p = p0[ip], rho = rho0[ip];
    pr.U[ip] = U; pr.V[ip] = V; pr.W[ip] = W; pr.p[ip] = p;
    r1.q[0][ip] = rho;
    r1.q[1][ip] = rho * U;
    r1.q[2][ip] = rho * V;
    r1.q[3][ip] = rho * W;
    r1.q[4][ip] = p / (g.gamma - 1.0) + 0.5 * rho * (U * U + V * V + W * W);
    const double T = solve_T(g, rho, p);
    pr.T[ip] = T;
    pr.mu[ip] = solve_mu(g, T, ip);
    pr.sos[ip] = solve_sos(g, rho, p);
}

// contravariant velocities, src/CurvilinearCSolver_Core.cpp:1190-1196; the flux
// kernel forms them itself from the same expression, so they are only
// materialised on request
__global__ void k_contravariant(Prim pr, Metrics m, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double U = ldnt(pr.U, ip), V = ldnt(pr.V, ip), W = ldnt(pr.W, ip);
    stnt(pr.Ucurv, ip, ldnt(m.J11, ip) * U + ldnt(m.J12, ip) * V + ldnt(m.J13, ip) * W);
    stnt(pr.Vcurv, ip, ldnt(m.J21, ip) * U + ldnt(m.J22, ip) * V + ldnt(m.J23, ip) * W);
    stnt(pr.Wcurv, ip, ldnt(m.J31, ip) * U + ldnt(m.J32, ip) * V + ldnt(m.J33, ip) * W);
}

// updateNonConservedData, src/CurvilinearCSolver_Core.cpp:1158-1198
__global__ void k_eos(Cons c, Prim pr, Gas g, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double rho = ldnt(c.q[0], ip);
    const double U = ldnt(c.q[1], ip) / rho, V = ldnt(c.q[2], ip) / rho, W = ldnt(c.q[3], ip) / rho;
    const double p = solve_p(g, rho, ldnt(c.q[4], ip), U, V, W);
    const double T = solve_T(g, rho, p);
    stnt(pr.U, ip, U); stnt(pr.V, ip, V); stnt(pr.W, ip, W);
    stnt(pr.p, ip, p);
    stnt(pr.T, ip, T);
    stnt(pr.mu, ip, solve_mu(g, T, ip));
    stnt(pr.sos, ip, solve_sos(g, rho, p));
}

// calcDtFromCFL wave-speed maximum, src/CurvilinearCSolver_Core.cpp:197-207
__global__ __launch_bounds__(kBlock) void k_cfl(Prim pr, Metrics m, double dx, double dy,
                                                double dz, double *partial, long n) {
    __shared__ double red[kBlock];
    double v = -100000.0;
    const long first = (long)blockIdx.y * n;  // this member's points
    for (long ip = first + (long)blockIdx.x * blockDim.x + threadIdx.x; ip < first + n;
         ip += (long)gridDim.x * blockDim.x) {
        const double s = ldnt(pr.sos, ip);
        const double u = ldnt(m.J11, ip) * (fabs(ldnt(pr.U, ip)) + s) / dx +
                         ldnt(m.J22, ip) * (fabs(ldnt(pr.V, ip)) + s) / dy +
                         ldnt(m.J33, ip) * (fabs(ldnt(pr.W, ip)) + s) / dz;
        v = u > v ? u : v;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// one workgroup per member
__global__ __launch_bounds__(kBlock) void k_cfl_finish(const double *partial, int nb,
                                                       StepScalars *scs, int const_cfl,
                                                       double cfl_target, int first_step,
                                                       double time0) {
    __shared__ double red[kBlock];
    StepScalars *sc = scs + blockIdx.x;
    partial += (long)blockIdx.x * nb;
    double v = -100000.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v = fmax(v, partial[i]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mx = red[0];
        sc->maxspeed = mx;
        if (const_cfl) sc->dt = cfl_target / mx;
        else sc->cfl = sc->dt * mx;
        if (first_step) sc->time = time0;
        else sc->time = sc->time + sc->dt;
    }
}

// one workgroup per member: out[2 b] = dt = cfl / maxspeed, out[2 b + 1] = maxspeed
__global__ __launch_bounds__(kBlock) void k_cfl_out(const double *partial, int nb, double cfl, double *out) {
    __shared__ double red[kBlock];
    partial += (long)blockIdx.x * nb;
    double v = -100000.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v = fmax(v, partial[i]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = cfl / red[0];
        out[2 * blockIdx.x + 1] = red[0];
    }
}

// Stress, heat flux and contravariant fluxes / J, src/CurvilinearCSolver_Core.cpp:548-561,615-752
__global__ __launch_bounds__(kBlock) void k_flux(FluxArgs a, Gas g, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const Metrics &m = a.m;
    const double J11 = ldnt(m.J11, ip), J12 = ldnt(m.J12, ip), J13 = ldnt(m.J13, ip);
    const double J21 = ldnt(m.J21, ip), J22 = ldnt(m.J22, ip), J23 = ldnt(m.J23, ip);
    const double J31 = ldnt(m.J31, ip), J32 = ldnt(m.J32, ip), J33 = ldnt(m.J33, ip);
    // velocity gradient tensor, gradU[i][j] = d u_i / d x_j
    const double dU1 = ldnt(a.dU[0], ip), dU2 = ldnt(a.dU[1], ip), dU3 = ldnt(a.dU[2], ip);
    const double dV1 = ldnt(a.dV[0], ip), dV2 = ldnt(a.dV[1], ip), dV3 = ldnt(a.dV[2], ip);
    const double dW1 = ldnt(a.dW[0], ip), dW2 = ldnt(a.dW[1], ip), dW3 = ldnt(a.dW[2], ip);
    const double dUdx = J11 * dU1 + J21 * dU2 + J31 * dU3;
    const double dUdy = J12 * dU1 + J22 * dU2 + J32 * dU3;
    const double dUdz = J13 * dU1 + J23 * dU2 + J33 * dU3;
    const double dVdx = J11 * dV1 + J21 * dV2 + J31 * dV3;
    const double dVdy = J12 * dV1 + J22 * dV2 + J32 * dV3;
    const double dVdz = J13 * dV1 + J23 * dV2 + J33 * dV3;
    const double dWdx = J11 * dW1 + J21 * dW2 + J31 * dW3;
    const double dWdy = J12 * dW1 + J22 * dW2 + J32 * dW3;
    const double dWdz = J13 * dW1 + J23 * dW2 + J33 * dW3;
    double T11 = (4.0 / 3.0) * dUdx - (2.0 / 3.0) * dVdy - (2.0 / 3.0) * dWdz;
    double T22 = -(2.0 / 3.0) * dUdx + (4.0 / 3.0) * dVdy - (2.0 / 3.0) * dWdz;
    double T33 = -(2.0 / 3.0) * dUdx - (2.0 / 3.0) * dVdy + (4.0 / 3.0) * dWdz;
    double T12 = dUdy + dVdx, T13 = dUdz + dWdx, T23 = dVdz + dWdy;
    const double mu = ldnt(a.mu, ip);
    double mu_eff, k_eff;
    if (a.les) {
        mu_eff = mu + ldnt(a.mu_sgs, ip);
        if (a.les == 2) k_eff = g.cp * mu / g.Pr + ldnt(a.k_sgs, ip);
        else k_eff = g.cp * (mu / g.Pr + ldnt(a.mu_sgs, ip) / a.Pr_t);
        if (mu_eff < 0.0) mu_eff = 0.0;
        if (k_eff < 0.0) k_eff = 0.0;
    } else {
        mu_eff = mu;
        k_eff = (g.cp / g.Pr) * mu;
    }
    T11 *= mu_eff; T22 *= mu_eff; T33 *= mu_eff;
    T12 *= mu_eff; T13 *= mu_eff; T23 *= mu_eff;
    const double T21 = T12, T31 = T13, T32 = T23;
    const double dT1 = ldnt(a.dT[0], ip), dT2 = ldnt(a.dT[1], ip), dT3 = ldnt(a.dT[2], ip);
    const double q1 = k_eff * (J11 * dT1 + J21 * dT2 + J31 * dT3);
    const double q2 = k_eff * (J12 * dT1 + J22 * dT2 + J32 * dT3);
    const double q3 = k_eff * (J13 * dT1 + J23 * dT2 + J33 * dT3);
    const double U = ldnt(a.U, ip), V = ldnt(a.V, ip), W = ldnt(a.W, ip), p = ldnt(a.p, ip);
    const double b1 = U * T11 + V * T12 + W * T13 + q1;
    const double b2 = U * T21 + V * T22 + W * T23 + q2;
    const double b3 = U * T31 + V * T32 + W * T33 + q3;
    const double iJ = 1.0 / ldnt(m.J, ip);
    // contravariant velocities (as updateNonConservedData forms them, :1190-1196)
    const double Uc = J11 * U + J12 * V + J13 * W;
    const double Vc = J21 * U + J22 * V + J23 * W;
    const double Wc = J31 * U + J32 * V + J33 * W;
    const double rho = ldnt(a.rho, ip), rhoU = ldnt(a.rhoU, ip), rhoV = ldnt(a.rhoV, ip), rhoW = ldnt(a.rhoW, ip);
    const double rhoE = ldnt(a.rhoE, ip);
    stnt(a.pre[0], ip, iJ * (rho * Uc));
    stnt(a.pre[5], ip, iJ * (rho * Vc));
    stnt(a.pre[10], ip, iJ * (rho * Wc));
    double F1, F2, F3, G1, G2, G3;
    F1 = rhoU * Uc + J11 * p; F2 = rhoU * Vc + J21 * p; F3 = rhoU * Wc + J31 * p;
    G1 = J11 * T11 + J12 * T21 + J13 * T31;
    G2 = J21 * T11 + J22 * T21 + J23 * T31;
    G3 = J31 * T11 + J32 * T21 + J33 * T31;
    stnt(a.pre[1], ip, iJ * (F1 - G1)); stnt(a.pre[6], ip, iJ * (F2 - G2)); stnt(a.pre[11], ip, iJ * (F3 - G3));
    F1 = rhoV * Uc + J12 * p; F2 = rhoV * Vc + J22 * p; F3 = rhoV * Wc + J32 * p;
    G1 = J11 * T12 + J12 * T22 + J13 * T32;
    G2 = J21 * T12 + J22 * T22 + J23 * T32;
    G3 = J31 * T12 + J32 * T22 + J33 * T32;
    stnt(a.pre[2], ip, iJ * (F1 - G1)); stnt(a.pre[7], ip, iJ * (F2 - G2)); stnt(a.pre[12], ip, iJ * (F3 - G3));
    F1 = rhoW * Uc + J13 * p; F2 = rhoW * Vc + J23 * p; F3 = rhoW * Wc + J33 * p;
    G1 = J11 * T13 + J12 * T23 + J13 * T33;
    G2 = J21 * T13 + J22 * T23 + J23 * T33;
    G3 = J31 * T13 + J32 * T23 + J33 * T33;
    stnt(a.pre[3], ip, iJ * (F1 - G1)); stnt(a.pre[8], ip, iJ * (F2 - G2)); stnt(a.pre[13], ip, iJ * (F3 - G3));
    F1 = (rhoE + p) * Uc; F2 = (rhoE + p) * Vc; F3 = (rhoE + p) * Wc;
    G1 = J11 * b1 + J12 * b2 + J13 * b3;
    G2 = J21 * b1 + J22 * b2 + J23 * b3;
    G3 = J31 * b1 + J32 * b2 + J33 * b3;
    stnt(a.pre[4], ip, iJ * (F1 - G1)); stnt(a.pre[9], ip, iJ * (F2 - G2)); stnt(a.pre[14], ip, iJ * (F3 - G3));
}

// one conserved variable: k2 and the register update of stage op; returns the
// value written to the stage register rk (stages that write it)
__device__ __forceinline__ double rk_update(const RkArgs &a, long ip, int q, double k2) {
    double *r1 = a.r1.q[q], *rk = a.rk.q[q], *r2 = a.r2.q[q];
    double n = 0.0;
    switch (a.op) {
        case kTvd1: n = ldnt(r1, ip) + a.c * k2; stnt(rk, ip, n); break;                         // TVDRK3.hpp:62-80
        case kTvd2: n = a.a * ldnt(r1, ip) + a.b * ldnt(rk, ip) + a.c * k2; stnt(rk, ip, n); break;  // :82-102
        case kTvd3: stnt(r2, ip, a.a * ldnt(r1, ip) + a.b * ldnt(rk, ip) + a.c * k2); break;         // :104-124
        case kRk4First: {  // RK4.hpp:53-70
            const double v = ldnt(r1, ip);
            stnt(r2, ip, v + a.a * k2); n = v + a.b * k2; stnt(rk, ip, n); break;
        }
        case kRk4Mid: stnt(r2, ip, ldnt(r2, ip) + a.a * k2); n = ldnt(r1, ip) + a.b * k2; stnt(rk, ip, n); break;  // :72-106
        case kRk4Last: stnt(r2, ip, ldnt(r2, ip) + a.a * k2); break;                                // :108-118
        case kLsFirst: n = ldnt(r1, ip) + a.b * k2; stnt(rk, ip, n); stnt(r1, ip, k2); break;      // KenRK4.hpp:58-74
        case kLsMid: { double v = k2 + a.a * ldnt(r1, ip); stnt(r1, ip, v); n = ldnt(rk, ip) + a.b * v; stnt(rk, ip, n); break; }  // :76-127
        case kLsLast: { double v = k2 + a.a * ldnt(r1, ip); stnt(r1, ip, v); stnt(r2, ip, ldnt(rk, ip) + a.b * v); break; }  // :128-145
    }
    return n;
}

// solveContinuity .. solveEnergy (src/CurvilinearCSolver_Core.cpp:790-947) fused
// with updateConservedData.  Momentum increments accumulate into the previous
// k2 before scaling (:839-840,872-873,904-905).  With a.eos set (a stage that
// writes rk and is not followed by a filter) updateNonConservedData of the new
// stage state follows in the same pass, from the values just computed.
__global__ __launch_bounds__(kBlock) void k_rhs_rk(RkArgs a, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double J = ldnt(a.J, ip);
    const StepScalars *sc = a.member_pts ? a.sc + (unsigned)ip / a.member_pts : a.sc;
    const double dtJ = sc->dt * J;
    double z[5];
    if (a.bc.sponge) {  // calcSpongeSource = sigma (avg - phi), src/CurvilinearCSolver.hpp:293-295
        const double sg = ldnt(a.bc.sigma, ip);
#pragma unroll
        for (int q = 0; q < 5; ++q) z[q] = (sg * (ldnt(a.bc.avg[q], ip) - ldnt(a.bc.qs.q[q], ip)) + 0.0) / J;
    } else {
        const double z0 = (0.0 + 0.0) / J;
#pragma unroll
        for (int q = 0; q < 5; ++q) z[q] = z0;
    }
    double k2[5];
    k2[0] = dtJ * (-ldnt(a.div[0], ip) + z[0]);
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        double v = ldnt(a.k2.q[q], ip) + (-ldnt(a.div[q], ip) + z[q]);
        k2[q] = v * dtJ;
    }
    {
        double v = -ldnt(a.div[4], ip) + z[4];
        k2[4] = v * dtJ;
    }
    if (a.bc.active) {  // postStepBCHandling, src/CurvilinearCSolver_BC.cpp:702-842
        const long NxNy = (long)a.bc.Nx * a.bc.Ny;
        const int i = (int)(ip % a.bc.Nx), j = (int)((ip / a.bc.Nx) % a.bc.Ny), k = (int)(ip / NxNy);
        const bool on[6] = {i == 0, i == a.bc.Nx - 1, j == 0, j == a.bc.Ny - 1, k == 0, k == a.bc.Nz - 1};
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (on[f] && a.bc.role[f] == kFaceWall) {
                k2[0] = -sc->dt * J * ldnt(a.bc.cont[f / 2], ip);
                k2[1] = 0.0; k2[2] = 0.0; k2[3] = 0.0; k2[4] = 0.0;
            }
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (on[f] && a.bc.role[f] == kFaceSponge) {
                k2[0] = 0.0; k2[1] = 0.0; k2[2] = 0.0; k2[3] = 0.0; k2[4] = 0.0;
            }
    }
    double nq[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        stnt(a.k2.q[q], ip, k2[q]);
        nq[q] = rk_update(a, ip, q, k2[q]);
    }
    if (a.eos) {  // as k_eos
        const Gas &g = a.g;
        const Prim &pr = a.pr;
        const double rho = nq[0];
        const double U = nq[1] / rho, V = nq[2] / rho, W = nq[3] / rho;
        const double p = solve_p(g, rho, nq[4], U, V, W);
        const double T = solve_T(g, rho, p);
        stnt(pr.U, ip, U); stnt(pr.V, ip, V); stnt(pr.W, ip, W);
        stnt(pr.p, ip, p);
        stnt(pr.T, ip, T);
        stnt(pr.mu, ip, solve_mu(g, T, ip));
        stnt(pr.sos, ip, solve_sos(g, rho, p));
    }
}

// Vreman eddy viscosity, src/VremanSGS.hpp:43-92 (abeta as the reference sums it)
__global__ __launch_bounds__(kBlock) void k_vreman(const double *dU0, const double *dU1,
                                                   const double *dU2, const double *dV0,
                                                   const double *dV1, const double *dV2,
                                                   const double *dW0, const double *dW1,
                                                   const double *dW2, Metrics m,
                                                   const double *__restrict__ rho, double *mu_sgs,
                                                   double dx, double dy, double dz, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= n) return;
    const double u1 = ldnt(dU0, ip), u2 = ldnt(dU1, ip), u3 = ldnt(dU2, ip);
    const double v1 = ldnt(dV0, ip), v2 = ldnt(dV1, ip), v3 = ldnt(dV2, ip);
    const double w1 = ldnt(dW0, ip), w2 = ldnt(dW1, ip), w3 = ldnt(dW2, ip);
    double d1v1, d2v1, d3v1, d1v2, d2v2, d3v2, d1v3, d2v3, d3v3;
    if (m.J11) {  // gradU = J^T d(u)/d(xi), src/CurvilinearCSolver_Core.cpp:548-561
        const double J11 = ldnt(m.J11, ip), J12 = ldnt(m.J12, ip), J13 = ldnt(m.J13, ip);
        const double J21 = ldnt(m.J21, ip), J22 = ldnt(m.J22, ip), J23 = ldnt(m.J23, ip);
        const double J31 = ldnt(m.J31, ip), J32 = ldnt(m.J32, ip), J33 = ldnt(m.J33, ip);
        d1v1 = J11 * u1 + J21 * u2 + J31 * u3;
        d2v1 = J12 * u1 + J22 * u2 + J32 * u3;
        d3v1 = J13 * u1 + J23 * u2 + J33 * u3;
        d1v2 = J11 * v1 + J21 * v2 + J31 * v3;
        d2v2 = J12 * v1 + J22 * v2 + J32 * v3;
        d3v2 = J13 * v1 + J23 * v2 + J33 * v3;
        d1v3 = J11 * w1 + J21 * w2 + J31 * w3;
        d2v3 = J12 * w1 + J22 * w2 + J32 * w3;
        d3v3 = J13 * w1 + J23 * w2 + J33 * w3;
    } else {  // the inputs are the Cartesian gradient, gradU[i][j] = d u_i / d x_j
        d1v1 = u1; d2v1 = u2; d3v1 = u3;
        d1v2 = v1; d2v2 = v2; d3v2 = v3;
        d1v3 = w1; d2v3 = w2; d3v3 = w3;
    }
    const double vol = dx * dy * dz / ldnt(m.J, ip);
    const double del = pow(vol, 1.0 / 3.0);
    const double d1 = del * del, d2 = d1, d3 = d1;
    const double b11 = d1 * d1v1 * d1v1 + d2 * d2v1 * d2v1 + d3 * d3v1 * d3v1;
    const double b12 = d1 * d1v1 * d1v2 + d2 * d2v1 * d2v2 + d3 * d3v1 * d3v2;
    const double b13 = d1 * d1v1 * d1v3 + d2 * d2v1 * d2v3 + d3 * d3v1 * d3v3;
    const double b22 = d1 * d1v2 * d1v2 + d2 * d2v2 * d2v2 + d3 * d3v2 * d3v2;
    const double b23 = d1 * d1v2 * d1v3 + d2 * d2v2 * d2v3 + d3 * d3v2 * d3v3;
    const double b33 = d1 * d1v3 * d1v3 + d2 * d2v3 * d2v3 + d3 * d3v3 * d3v3;
    const double abeta = d1v1 * d1v1 + d1v2 * d1v2 + d1v3 * d1v3 + d2v1 * d2v1 + d2v2 * d2v2 +
                         d2v3 * d2v3 + d3v1 * d3v1 + d2v3 * d2v3 + d3v3 * d3v3;
    double bbeta = b11 * b22 - (b12 * b12) + b11 * b33 - (b13 * b13) + b22 * b33 - (b23 * b23);
    const double eps = 1e-12;
    if (bbeta < eps) bbeta = eps;
    const double ss = sqrt(bbeta / (abeta + eps));
    const double c = 2.5 * 0.17 * 0.17;
    const double ms = c * ldnt(rho, ip) * ss;
    const double cap = 0.05 * d1 * sqrt(2.0 * abeta);
    stnt(mu_sgs, ip, ms < cap ? ms : cap);
}

__global__ void k_fill(double *p, double v, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip < n) p[ip] = v;
}

// src/AbstractSGS.cpp:15-131: each term divided by the line length, summed in
// index order, then broadcast back along the line
template <int DIR>
__device__ __forceinline__ void xi_average_line(const double *phi, double *phiF, int Nx, int Ny, int Nz,
                                                double *line = nullptr) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = DIR == 0 ? Nx : DIR == 1 ? Ny : Nz;
    const long nlines = (long)Nx * Ny * Nz / nl;
    if (q >= nlines) return;
    long base = 0, stride = 1;
    if (DIR == 0) { base = q * Nx; stride = 1; }
    if (DIR == 1) { base = (q / Nx) * (long)Nx * Ny + q % Nx; stride = Nx; }
    if (DIR == 2) { base = q; stride = (long)Nx * Ny; }
    double acc = 0.0;
    for (int i = 0; i < nl; ++i) acc += phi[base + i * stride] / (double)nl;
    if (line) {
        line[q] = acc;
        return;
    }
    for (int i = 0; i < nl; ++i) phiF[base + i * stride] = acc;
}

template <int DIR>
__global__ void k_xi_average(const double *phi, double *phiF, int Nx, int Ny, int Nz) {
    xi_average_line<DIR>(phi, phiF, Nx, Ny, Nz);
}

struct XiFields {
    const double *phi[8];
    double *phiF[8];
    double *line[8];  // per-line means instead of phiF (null: write phiF)
};

// x-lines (contiguous): a workgroup stages L whole lines through LDS with
// coalesced loads, one thread per line sums f / N in line order (the same
// rounding as xi_average_line), and the mean is written back coalesced
constexpr int kXiThreads = 256;
inline int xi1_lines(int Nx) { return Nx >= 8192 ? 1 : (8192 / Nx < 32 ? 8192 / Nx : 32); }

__global__ __launch_bounds__(kXiThreads) void k_xi1_average_tiled(XiFields f, int Nx, long nlines, int L) {
    extern __shared__ double tile[];  // [L][Nx + 1]
    const double *phi = f.phi[blockIdx.y];
    double *phiF = f.phiF[blockIdx.y];
    const long q0 = (long)blockIdx.x * L;
    const int nl = (int)(nlines - q0 < L ? nlines - q0 : L);
    const int S = Nx + 1;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int l = e / Nx, i = e - l * Nx;
        tile[l * S + i] = phi[q0 * Nx + e];
    }
    __syncthreads();
    double *line = f.line[blockIdx.y];
    if ((int)threadIdx.x < nl) {
        double acc = 0.0;
        const double *row = tile + threadIdx.x * S;
        for (int i = 0; i < Nx; ++i) acc += row[i] / (double)Nx;
        tile[threadIdx.x * S + Nx] = acc;  // the padding column holds the mean
        if (line) line[q0 + threadIdx.x] = acc;
    }
    if (line) return;  // uniform: the whole workgroup leaves
    __syncthreads();
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int l = e / Nx;
        phiF[q0 * Nx + e] = tile[l * S + Nx];
    }
}

template <int DIR>
__global__ void k_xi_average_fields(XiFields f, int Nx, int Ny, int Nz) {
    xi_average_line<DIR>(f.phi[blockIdx.y], f.phiF[blockIdx.y], Nx, Ny, Nz, f.line[blockIdx.y]);
}

__global__ __launch_bounds__(kBlock) void k_sum_partial(const double *phi, double *partial, long n) {
    __shared__ double red[kBlock];
    double v = 0.0;
    for (long ip = (long)blockIdx.x * blockDim.x + threadIdx.x; ip < n; ip += (long)gridDim.x * blockDim.x)
        v += phi[ip];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// per-block minimum and maximum of phi into partial[2b], partial[2b+1]
__global__ __launch_bounds__(kBlock) void k_minmax_partial(const double *phi, double *partial, long n) {
    __shared__ double lo[kBlock], hi[kBlock];
    double a = 1000000.0, b = -1000000.0;  // getRange's fold seeds, src/Utils.cpp:417,421
    for (long ip = (long)blockIdx.x * blockDim.x + threadIdx.x; ip < n; ip += (long)gridDim.x * blockDim.x) {
        const double v = phi[ip];
        a = fmin(v, a);
        b = fmax(v, b);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = lo[0];
        partial[2 * blockIdx.x + 1] = hi[0];
    }
}

// mean of the block partial sums, written to *mean
__global__ __launch_bounds__(kBlock) void k_mean_finish(const double *partial, int nb, double *mean, long n) {
    __shared__ double red[kBlock];
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v += partial[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *mean = red[0] / (double)n;
}

__global__ void k_fill_from(double *p, const double *v, long n) {
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip < n) p[ip] = *v;
}

}  // namespace

hipError_t launch_init_state(const double *rho0, const double *U0, const double *V0,
                             const double *W0, const double *p0, Cons r1, Prim pr, Metrics m,
                             Gas g, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_init_state, grid_for(n), dim3(kBlock), 0, st, rho0, U0, V0, W0, p0, r1, pr,
                       m, g, n);
    return hipGetLastError();
}

hipError_t launch_eos(Cons in, Prim pr, Gas g, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_eos, grid_for(n), dim3(kBlock), 0, st, in, pr, g, n);
    return hipGetLastError();
}

hipError_t launch_contravariant(Prim pr, Metrics m, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_contravariant, grid_for(n), dim3(kBlock), 0, st, pr, m, n);
    return hipGetLastError();
}

int cfl_blocks(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)(b < 2048 ? b : 2048);
}

hipError_t launch_cfl(Prim pr, Metrics m, double dx, double dy, double dz, double *partial, long n,
                      int nmember, hipStream_t st) {
    if (nmember < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cfl, dim3(cfl_blocks(n), nmember), dim3(kBlock), 0, st, pr, m, dx, dy, dz, partial, n);
    return hipGetLastError();
}

hipError_t launch_cfl_finish(const double *partial, int nblocks, StepScalars *sc, int nmember, int const_cfl,
                             double cfl_target, int first_step, double time0, hipStream_t st) {
    if (nmember < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cfl_finish, dim3(nmember), dim3(kBlock), 0, st, partial, nblocks, sc, const_cfl,
                       cfl_target, first_step, time0);
    return hipGetLastError();
}

hipError_t launch_cfl_out(const double *partial, int nblocks, int nmember, double cfl, double *out,
                          hipStream_t st) {
    if (nmember < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cfl_out, dim3(nmember), dim3(kBlock), 0, st, partial, nblocks, cfl, out);
    return hipGetLastError();
}

hipError_t launch_flux(const FluxArgs &a, Gas g, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_flux, grid_for(n), dim3(kBlock), 0, st, a, g, n);
    return hipGetLastError();
}

hipError_t launch_rhs_rk(const RkArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_rhs_rk, grid_for(n), dim3(kBlock), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_vreman(const double *const dU[3], const double *const dV[3],
                         const double *const dW[3], Metrics m, const double *rho, double *mu_sgs,
                         double dx, double dy, double dz, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_vreman, grid_for(n), dim3(kBlock), 0, st, dU[0], dU[1], dU[2], dV[0], dV[1],
                       dV[2], dW[0], dW[1], dW[2], m, rho, mu_sgs, dx, dy, dz, n);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, double v, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_xi_average(const double *phi, double *phiF, int dir, int Nx, int Ny, int Nz,
                             hipStream_t st) {
    if (dir < 0 || dir > 2) return hipErrorInvalidValue;
    const int nl = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
    const long nlines = (long)Nx * Ny * Nz / nl;
    if (dir == 0) {
        const double *p1[1] = {phi};
        double *o1[1] = {phiF};
        return launch_xi_average_fields(p1, o1, 1, 0, Nx, Ny, Nz, st);
    }
    if (dir == 1) hipLaunchKernelGGL(k_xi_average<1>, grid_for(nlines), dim3(kBlock), 0, st, phi, phiF, Nx, Ny, Nz);
    if (dir == 2) hipLaunchKernelGGL(k_xi_average<2>, grid_for(nlines), dim3(kBlock), 0, st, phi, phiF, Nx, Ny, Nz);
    return hipGetLastError();
}

hipError_t launch_xi_average_fields(const double *const *phi, double *const *phiF, int nf, int dir, int Nx,
                                    int Ny, int Nz, hipStream_t st, double *const *line) {
    if (nf < 1 || nf > 8 || dir < 0 || dir > 2) return hipErrorInvalidValue;
    XiFields f{};
    for (int k = 0; k < nf; ++k) {
        f.phi[k] = phi[k];
        f.phiF[k] = phiF ? phiF[k] : nullptr;
        f.line[k] = line ? line[k] : nullptr;
    }
    const long nlines = (long)Nx * Ny * Nz / (dir == 0 ? Nx : dir == 1 ? Ny : Nz);
    const dim3 grid((unsigned)((nlines + kBlock - 1) / kBlock), nf);
    if (dir == 0) {
        const int L = xi1_lines(Nx);
        const size_t lds = sizeof(double) * (size_t)L * (Nx + 1);
        hipLaunchKernelGGL(k_xi1_average_tiled, dim3((unsigned)((nlines + L - 1) / L), nf), dim3(kXiThreads), lds, st,
                           f, Nx, nlines, L);
        return hipGetLastError();
    }
    if (dir == 1) hipLaunchKernelGGL(k_xi_average_fields<1>, grid, dim3(kBlock), 0, st, f, Nx, Ny, Nz);
    if (dir == 2) hipLaunchKernelGGL(k_xi_average_fields<2>, grid, dim3(kBlock), 0, st, f, Nx, Ny, Nz);
    return hipGetLastError();
}

hipError_t launch_minmax_partial(const double *phi, double *partial, int nblocks, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_minmax_partial, dim3(nblocks), dim3(kBlock), 0, st, phi, partial, n);
    return hipGetLastError();
}

hipError_t launch_global_mean(const double *phi, double *phiF, double *partial, long n,
                              hipStream_t st) {
    const int nb = cfl_blocks(n);
    hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(kBlock), 0, st, phi, partial, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the mean goes to partial[nb] (the buffer holds 2 * 2048 doubles)
    hipLaunchKernelGGL(k_mean_finish, dim3(1), dim3(kBlock), 0, st, partial, nb, partial + nb, n);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill_from, grid_for(n), dim3(kBlock), 0, st, phiF, partial + nb, n);
    return hipGetLastError();
}

}  // namespace mc3d
