// Pointwise kernels of the dynamic Smagorinsky model (DSM), launch interface.
//
// The model (src/DSMSGS.hpp:124-444) alternates pointwise products with
// three-direction test filters; the filters are line operators launched by
// the host class, these kernels are the pointwise stages between them.
#pragma once

#include <hip/hip_runtime.h>

#include "pointwise.h"

namespace mc3d {

// strain rate, |S| and the grid-level M_ij terms (src/DSMSGS.hpp:127-175)
struct DsmStrainArgs {
    const double *dU[3], *dV[3], *dW[3];  // computational-space velocity derivatives
    Metrics m;
    const double *rho;
    double *S[6];     // S00 S01 S02 S11 S12 S22
    double *Smag;
    double *M2[6];    // M00_2 M01_2 M02_2 M11_2 M12_2 M22_2
};

// M_ij -= test-level term; L_ij tensors before filtering (src/DSMSGS.hpp:200-260)
struct DsmMLArgs {
    double *M[6];            // in/out: filtered grid-level M_ij
    const double *S_hat[6];
    const double *rho_hat, *Smag_hat;
    const double *rho, *rhoU, *rhoV, *rhoW;
    double *Lt[6];           // out: rhoU_i rhoU_j / rho
};

// L_ij -= hat products; CIdenom = 2 rho |S|^2 (src/DSMSGS.hpp:262-300)
struct DsmLArgs {
    double *L[6];
    const double *rhoU_hat, *rhoV_hat, *rhoW_hat, *rho_hat;
    const double *rho, *Smag;
    double *CIdenom;  // null when tau_kk is not used
};

// contractions Lkk, L_ij M_ij, M_ij M_ij (and CIdenom) (src/DSMSGS.hpp:300-318)
struct DsmContractArgs {
    const double *L[6], *M[6];
    const double *rho_hat, *Smag_hat, *alpha_hat;  // alpha_hat null when tau_kk is not used
    double *CIdenom;
    double *Lkk, *LijMij, *MijMij;
};

// mu_sgs, tau_kk and the coefficient fields (src/DSMSGS.hpp:319-341)
struct DsmMuArgs {
    const double *rho, *Smag, *LijMij_avg, *MijMij_avg, *Lkk_avg, *CIdenom_avg;
    const double *J;
    double *mu_sgs, *taukk, *C, *CI;  // C, CI null unless the coefficient range is kept
    int use_taukk;
    double dx, dy, dz;
};

// N_i, K_i, rho T and the Cartesian temperature gradient (src/DSMSGS.hpp:345-380)
struct DsmNKArgs {
    const double *dT[3];
    Metrics m;
    const double *rho, *Smag, *rhoU, *rhoV, *rhoW, *T;
    double *N[3], *K[3], *rhoT, *gT[3];
};

// N_i, K_i at the test level and their contractions (src/DSMSGS.hpp:382-420)
struct DsmNNArgs {
    const double *rho_hat, *Smag_hat, *g_hat[3], *NRHS[3];
    const double *rhoU_hat, *rhoV_hat, *rhoW_hat, *rhoT_hat, *rho, *KRHS[3];
    double *NiNi, *NiKi;
};

// k_sgs and Pr_t (src/DSMSGS.hpp:421-444)
struct DsmKArgs {
    const double *NiNi_avg, *NiKi_avg, *rho, *Smag, *C, *J;
    double *k_sgs, *Prt;  // Prt null unless the coefficient range is kept
    double cp, dx, dy, dz;
};

hipError_t launch_dsm_strain(const DsmStrainArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_ml(const DsmMLArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_l(const DsmLArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_contract(const DsmContractArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_mu(const DsmMuArgs &a, long n, hipStream_t st);
// tau_kk folded into the pressure before the temperature gradient
// (src/CurvilinearCSolver_Core.cpp:571-584): p -= tau_kk / 2, then T, mu, sos
hipError_t launch_dsm_taukk_eos(const double *rho, const double *taukk, Prim pr, Gas g, long n,
                                hipStream_t st);
hipError_t launch_dsm_nk(const DsmNKArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_nn(const DsmNNArgs &a, long n, hipStream_t st);
hipError_t launch_dsm_ksgs(const DsmKArgs &a, long n, hipStream_t st);

}  // namespace mc3d
