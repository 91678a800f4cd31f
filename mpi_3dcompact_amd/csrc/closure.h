// Learned SGS closure: a 7-64-64-64-2 MLP (tanh-form GELU) replacing the
// dynamic procedure of the compressible DSM (src/DSMSGS.hpp:124-444).
//
// Numerics: every matrix product takes BF16 operands (round to nearest even)
// on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) and accumulates in
// FP32; biases, GELU, loss and Adam are FP32; the per-point features are
// formed in FP64 and rounded to FP32 before standardisation.
#pragma once

#include <hip/hip_runtime.h>

#include "pointwise.h"

namespace mc3d {

constexpr int kMlpIn = 7, kMlpHid = 64, kMlpOut = 2;
// flat FP32 parameter layout: W1[64][7] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2]
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHid * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid * kMlpHid;
constexpr int kMlpOffW4 = kMlpOffB3 + kMlpHid;
constexpr int kMlpOffB4 = kMlpOffW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kMlpOffB4 + kMlpOut;  // 8962
constexpr int kMlpTile = 64;                      // samples per training workgroup

// standardisation and the trusted input range of a trained network
struct MlpScaling {
    float in_mean[kMlpIn], in_std[kMlpIn];
    float out_mean[kMlpOut], out_std[kMlpOut];
    float lo[kMlpIn], hi[kMlpIn];  // raw-feature range; outside it the point falls back
};

// Gradient of L = sum((y - t)^2) / (n_total * 2) over samples idx[0..n)
// (idx null: 0..n), standardised inputs X[n_all][7] and targets T[n_all][2].
// partial: [nblocks][kMlpParams] gradients, lossp: [nblocks] sums of squares,
// nblocks = ceil(n / 64).
hipError_t launch_mlp_grad_partials(const float *params, const float *X, const float *T, const int *idx,
                                    long n, long n_total, float *partial, float *lossp, hipStream_t st);
// grad[p] = sum_b partial[b][p] (b ascending); loss[0] = sum_b lossp[b] / (n_total * 2)
hipError_t launch_mlp_reduce(const float *partial, const float *lossp, int nblocks, long n_total, float *grad,
                             float *loss, hipStream_t st);
// Adam update of params (FP32) with moments m, v; c1 = 1 - b1^t, c2 = 1 - b2^t
struct AdamArgs {
    float lr, b1, b2, eps, c1, c2;
};
hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *grad, AdamArgs a, hipStream_t st);
// Y[n][2] = network(X[n][7]) on standardised values (the register-resident
// inference path the solver uses)
hipError_t launch_mlp_infer(const float *params, const float *X, float *Y, long n, hipStream_t st);

// The seven raw features of every point (FP32 values stored as FP64 fields),
// for averaging before the network: the closure then sees its inputs averaged
// as the DSM averages L_ij M_ij and M_ij M_ij (LESAVERAGING).
struct FeatureArgs {
    const double *dU[3], *dV[3], *dW[3];
    Metrics m;
    double vol;
    double *feat[kMlpIn];
    double *smag, *delta2;  // |S| and Delta^2 of each point (null: not written)
};
hipError_t launch_closure_features(const FeatureArgs &a, long n, hipStream_t st);

// Per-point closure on the solver's fields: Cartesian gradient from the
// computational derivatives and metrics (src/CurvilinearCSolver_Core.cpp:548-561),
// features, network, then
//     mu_sgs = rho C Delta^2 |S|,  k_sgs = cp (C / Pr_t) rho Delta^2 |S|,
// Delta = (dxi deta dzeta / J)^(1/3).  flag[p] = 1 where a raw feature lies
// outside [lo, hi]; such inputs are clamped to the range for the network (and
// mu_sgs, k_sgs are written there too); blockcount[b] = number of flagged
// points of workgroup b.
struct ClosureArgs {
    const double *dU[3], *dV[3], *dW[3];
    Metrics m;
    const double *rho;
    const float *params;
    MlpScaling sc;
    double vol;  // dxi * deta * dzeta
    double cp;
    double *mu_sgs, *k_sgs;
    unsigned char *flag;
    int *blockcount;
    const double *feat[kMlpIn];  // averaged raw features; null: formed at each point
    const double *smag, *delta2;  // with feat: |S| and Delta^2 of each point (FeatureArgs)
};
int closure_blocks(long n);  // workgroups launch_learned_sgs uses (size of blockcount)
hipError_t launch_learned_sgs(const ClosureArgs &a, long n, hipStream_t st);
// With LESAVERAGING XI1 / XI2 / XI3 the averaged inputs, and so C and C/Pr_t,
// are constant along each line: the network runs once per line on the line
// means of the features (feat[k][q], nlines lines), writing C[q], D[q] = C/Pr_t
// and the out-of-range flag of each line; blockcount[b] = flagged lines of
// workgroup b times the line length.  The same numerics as launch_learned_sgs.
struct LineClosureArgs {
    const double *feat[kMlpIn];
    const float *params;
    MlpScaling sc;
    long nlines;
    int line_len;
    float *C, *D;
    unsigned char *lflag;
    int *blockcount;
};
int line_closure_blocks(long nlines);  // workgroups launch_line_closure uses
hipError_t launch_line_closure(const LineClosureArgs &a, hipStream_t st);
// every point: mu_sgs = (rho Delta^2) |S| C[line], k_sgs = (cp rho Delta^2 |S|) D[line],
// flag = the line's flag (lines of direction dir numbered as xi_line_of numbers them)
struct ClosureApplyArgs {
    const double *rho, *smag, *delta2;
    const float *C, *D;
    const unsigned char *lflag;
    double cp;
    int dir, Nx, Ny;
    double *mu_sgs, *k_sgs;
    unsigned char *flag;
};
hipError_t launch_closure_apply(const ClosureApplyArgs &a, long n, hipStream_t st);
// total = sum of blockcount[0..nb) into count[0]
hipError_t launch_count_sum(const int *blockcount, int nb, int *count, hipStream_t st);
// where flag[p]: mu_sgs, k_sgs = the DSM's values at that point
struct FallbackArgs {
    const unsigned char *flag;
    const double *mu_dsm, *k_dsm;
    double *mu_sgs, *k_sgs;
};
hipError_t launch_closure_fallback(const FallbackArgs &a, long n, hipStream_t st);

// Training pairs from a DSM state: for i < count, point p = offset + i * stride:
// X[i][7] = raw FP32 features, T[i][0] = C, T[i][1] = C / Pr_t (the DSM's
// coefficient fields), from the same gradients the DSM used.
struct SampleArgs {
    const double *dU[3], *dV[3], *dW[3];
    Metrics m;
    const double *C, *Prt;
    double vol;
    long offset, stride, count;
    float *X, *T;
    const double *feat[kMlpIn];  // averaged raw features; null: formed at each point
};
hipError_t launch_closure_samples(const SampleArgs &a, hipStream_t st);

}  // namespace mc3d
