// Pointwise kernels of the solver step (launch interface).
#pragma once

#include <hip/hip_runtime.h>

namespace mc3d {

// Ideal gas, src/IdealGas.hpp:35-55
struct Gas {
    double gamma, Pr, R_gas, cp, T_ref, mu_ref;
    // members stacked along z with their own mu_ref: mu_member[ip / member_pts]
    // replaces mu_ref when mu_member is set
    const double *mu_member;
    unsigned member_pts;
};

// Streaming access of the pointwise kernels: each field value is touched once
// per launch and the fields exceed every cache, so loads and stores are
// non-temporal (measured: 5.5 -> 6.0 TB/s for kernels of 13-48 streams).
__device__ __forceinline__ double ldnt(const double *p, long i) { return __builtin_nontemporal_load(p + i); }
__device__ __forceinline__ void stnt(double *p, long i, double v) { __builtin_nontemporal_store(v, p + i); }

__device__ __forceinline__ double member_mu_ref(const Gas &g, long ip) {
    return g.mu_member ? g.mu_member[(unsigned)ip / g.member_pts] : g.mu_ref;
}

struct Metrics {
    const double *J, *J11, *J12, *J13, *J21, *J22, *J23, *J31, *J32, *J33;
};

struct Cons {  // one register set of conserved variables: rho, rhoU, rhoV, rhoW, rhoE
    double *q[5];
};

struct Prim {  // non-conserved data
    double *U, *V, *W, *p, *T, *mu, *sos, *Ucurv, *Vcurv, *Wcurv;
};

// time-step scalars kept on the device so whole steps can be captured
struct StepScalars {
    double dt;
    double cfl;
    double time;
    double maxspeed;
};

enum RkOp {
    kTvd1 = 0, kTvd2, kTvd3,          // TVDRK3 stages
    kRk4First, kRk4Mid, kRk4Last,     // classic RK4
    kLsFirst, kLsMid, kLsLast         // 2N low-storage (KenRK4, LSLDDRK4)
};

// face ids X0 X1 Y0 Y1 Z0 Z1 = 0..5; a face of direction d = f / 2, the high end when f % 2
enum FaceRole : signed char { kFaceNone = 0, kFaceWall = 1, kFaceSponge = 2 };

// Post-stage face handling and the sponge source, applied inside the RK stage
// kernel: walls keep only the normal continuity flux, k2_rho = -dt J cont_d
// (faces in the order X0..Z1), sponge faces keep nothing (:702-842); the
// sponge source sigma (avg - q) / J enters every equation (:790-947,
// src/CurvilinearCSolver.hpp:293-295).
struct RkBc {
    int active;                // any wall or sponge face
    signed char role[6];       // FaceRole per face
    int Nx, Ny, Nz;
    const double *cont[3];     // d/dxi_d of the direction-d continuity pre-flux (wall directions)
    int sponge;                // sponge source on
    const double *sigma;
    const double *avg[5];
    Cons qs;                   // stage state the sponge relaxes
};

struct RkArgs {
    Cons r1, rk, k2, r2;
    const double *div[5];  // summed flux divergences (cont, mom1..3, engy)
    const double *J;
    const StepScalars *sc;  // one per member when member_pts is set: sc[ip / member_pts]
    unsigned member_pts;
    int op;
    double a, b, c;        // stage coefficients
    int eos;               // also update the non-conserved data from the new rk state
    Prim pr;
    Gas g;
    RkBc bc;
};

struct FluxArgs {
    const double *dU[3], *dV[3], *dW[3], *dT[3];
    Metrics m;
    const double *U, *V, *W, *p, *mu, *mu_sgs, *k_sgs;
    const double *rho, *rhoU, *rhoV, *rhoW, *rhoE;
    double *pre[15];
    int les;  // 0 none, 1 Vreman-type (k from mu_sgs/Pr_t), 2 DSM (k_sgs)
    double Pr_t;
};

hipError_t launch_init_state(const double *rho0, const double *U0, const double *V0,
                             const double *W0, const double *p0, Cons r1, Prim pr, Metrics m,
                             Gas g, long n, hipStream_t st);
hipError_t launch_eos(Cons in, Prim pr, Gas g, long n, hipStream_t st);
// Ucurv, Vcurv, Wcurv from U, V, W and the metrics (on request only)
hipError_t launch_contravariant(Prim pr, Metrics m, long n, hipStream_t st);
// CFL: for each of nmember members of n points (stacked), partial maxima
// into partial[member * cfl_blocks(n) + block], then one finishing workgroup
// per member that sets its dt (const CFL) or cfl (const dt) and advances its time.
int cfl_blocks(long n);
hipError_t launch_cfl(Prim pr, Metrics m, double dx, double dy, double dz, double *partial,
                      long n, int nmember, hipStream_t st);
hipError_t launch_cfl_finish(const double *partial, int nblocks, StepScalars *sc, int nmember, int const_cfl,
                             double cfl_target, int first_step, double time0, hipStream_t st);
// the same maxima into out[2 b] = cfl / maxspeed, out[2 b + 1] = maxspeed of each member
hipError_t launch_cfl_out(const double *partial, int nblocks, int nmember, double cfl, double *out,
                          hipStream_t st);
hipError_t launch_flux(const FluxArgs &a, Gas g, long n, hipStream_t st);
hipError_t launch_rhs_rk(const RkArgs &a, long n, hipStream_t st);
hipError_t launch_vreman(const double *const dU[3], const double *const dV[3],
                         const double *const dW[3], Metrics m, const double *rho, double *mu_sgs,
                         double dx, double dy, double dz, long n, hipStream_t st);
hipError_t launch_fill(double *p, double v, long n, hipStream_t st);
// doAveraging XI1/XI2/XI3: average along direction dir (in place allowed)
hipError_t launch_xi_average(const double *phi, double *phiF, int dir, int Nx, int Ny, int Nz,
                             hipStream_t st);
// the same for nf <= 8 fields in one launch (phi[f] -> phiF[f], in place allowed);
// with line set, the means go to line[f][q] (one value per line q, numbered as
// xi_line_of numbers them) and phiF is not written
hipError_t launch_xi_average_fields(const double *const *phi, double *const *phiF, int nf, int dir, int Nx,
                                    int Ny, int Nz, hipStream_t st, double *const *line = nullptr);

// line number of point p for direction dir (lines of direction dir, as the
// line averages number them)
__device__ __forceinline__ long xi_line_of(long p, int dir, int Nx, int Ny) {
    if (dir == 0) return p / Nx;
    if (dir == 1) return (p / ((long)Nx * Ny)) * Nx + p % Nx;
    return p % ((long)Nx * Ny);
}
// per-block {min, max} pairs of phi into partial[2 * nblocks] (getRange)
hipError_t launch_minmax_partial(const double *phi, double *partial, int nblocks, long n, hipStream_t st);
// doAveraging GLOBAL: phiF = mean(phi); partial needs cfl_blocks(n) + 1 doubles
hipError_t launch_global_mean(const double *phi, double *phiF, double *partial, long n,
                              hipStream_t st);

}  // namespace mc3d
