// Metric terms from mesh coordinates: pointwise and wrap-value kernels
// (src/AbstractSingleBlockMesh.cpp:23-1439).
#pragma once

#include <hip/hip_runtime.h>

namespace mc3d {

// Wrap values of the transformed-periodic derivative along direction dir:
// for every line q (line order of the line operators) and slot
// {Nm3, Nm2, Nm1, Np1, Np2, Np3}, at the line's points N-3, N-2, N-1 (shift
// -T) and 0, 1, 2 (shift +T):
//     value = sum_t (A_t[p] -/+ T_t) * B_t[p]       (B_t null: no factor)
// summed over t < nterms in order, as the reference forms them.
struct WrapArgs {
    const double *A[3];
    const double *B[3];
    double T[3];
    int nterms;
    double *out[6];
};
hipError_t launch_tpb_wraps(const WrapArgs &a, int dir, int Nx, int Ny, int Nz, hipStream_t st);

// Va_j = xE1j y, Vb_j = xE1j z, Vc_j = xE2j z (src/AbstractSingleBlockMesh.cpp:298-310)
hipError_t launch_metric_products(const double *const xE1[3], const double *const xE2[3], const double *y,
                                  const double *z, double *const Va[3], double *const Vb[3],
                                  double *const Vc[3], long n, hipStream_t st);
// J_ij / J from the cross derivatives dV[j][k] = d V_j / d xi_k (:1011-1023)
// and pre_k = x J_k1 + y J_k2 + z J_k3 (:1069-1073)
struct CofactorArgs {
    const double *dVa[3][3], *dVb[3][3], *dVc[3][3];
    const double *x, *y, *z;
    double *Jm[3][3];
    double *pre[3];
};
hipError_t launch_metric_cofactors(const CofactorArgs &a, long n, hipStream_t st);
// J = 1 / ((1/3)(Jdet1 + Jdet2 + Jdet3)), J_ij *= J (:1296-1309)
hipError_t launch_metric_final(const double *const Jdet[3], double *J, double *const Jm[9], long n,
                               hipStream_t st);

}  // namespace mc3d
