// Metric terms from mesh coordinates on the GPU.
#include "mesh.hpp"

#include <algorithm>

#include "../metrics.h"

namespace mc3d {

// AbstractSingleBlockMesh::solveForJacobians, src/AbstractSingleBlockMesh.cpp:23-1309:
// derivatives of x and y, the products V = xE {y, z}, their cross
// derivatives, J_ij / J as their differences, J from the divergence of
// x J_i, then J_ij *= J.  Periodic directions use the transformed-periodic
// derivative with wrap values shifted by the direction's translation.  With
// the 7-point stencil the reference passes Nm2 in the Nm3 slot for the zeta
// derivative of y (:273) and the xi derivative of Vc3 (:706); kept.
// Directions are visited eta, xi, zeta, as the reference orders them.
void solve_for_jacobians(AbstractDerivatives *const dv[3], const MeshSpec &m, const double *x,
                         const double *y, const double *z, double *const out[10], hipStream_t st) {
    const long N = (long)m.N[0] * m.N[1] * m.N[2];
    // scratch: 6 xE + 9 V + 18 dV + 3 pre + 3 Jdet fields, 6 wrap planes
    const int nf = 6 + 9 + 18 + 3 + 3;
    long maxl = 0;
    for (int d = 0; d < 3; ++d) maxl = std::max(maxl, N / m.N[d]);
    void *mem = nullptr;
    hip_check(hipMalloc(&mem, sizeof(double) * ((size_t)N * nf + (size_t)6 * maxl)), "metric scratch");
    struct Free {
        void *p;
        hipStream_t s;
        ~Free() {
            (void)hipStreamSynchronize(s);  // also on an error path: queued work may use the scratch
            (void)hipFree(p);
        }
    } guard{mem, st};
    double *f = static_cast<double *>(mem);
    auto take = [&]() { double *r = f; f += N; return r; };
    double *xE[2][3], *Va[3], *Vb[3], *Vc[3], *dVa[3][3] = {}, *dVb[3][3] = {}, *dVc[3][3] = {};
    double *pre[3], *Jdet[3];
    for (auto &r : xE) for (auto &p : r) p = take();
    for (int j = 0; j < 3; ++j) { Va[j] = take(); Vb[j] = take(); Vc[j] = take(); }
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
            if (j != k) { dVa[j][k] = take(); dVb[j][k] = take(); dVc[j][k] = take(); }
    for (int k = 0; k < 3; ++k) { pre[k] = take(); Jdet[k] = take(); }
    double *w[6];
    for (auto &p : w) { p = f; f += maxl; }
    const double *xyz[3] = {x, y, z};

    // d f / d xi_d; periodic directions with wrap values of sum_t (A_t -/+ T_t) B_t
    auto deriv = [&](int d, const double *in, double *o, int nterms, const double *const *A,
                     const double *const *B, const double *tr, bool nm3_typo) {
        if (!m.periodic[d]) {
            dv[d]->calc1stDerivField(in, o);
            return;
        }
        WrapArgs wa{};
        for (int t = 0; t < nterms; ++t) { wa.A[t] = A[t]; wa.B[t] = B ? B[t] : nullptr; wa.T[t] = tr[t]; }
        wa.nterms = nterms;
        for (int s = 0; s < 6; ++s) wa.out[s] = w[s];
        hip_check(launch_tpb_wraps(wa, d, m.N[0], m.N[1], m.N[2], st), "metric wrap values");
        if (m.bw7[d])
            dv[d]->calc1stDerivField_TPB(in, o, nm3_typo ? w[1] : w[0], w[1], w[2], w[3], w[4], w[5]);
        else
            dv[d]->calc1stDerivField_TPB(in, o, w[1], w[2], w[3], w[4]);
    };
    // x and y derivatives, :40-281
    for (int d : {1, 0, 2})
        for (int i = 0; i < 2; ++i) {
            const double *A[1] = {xyz[i]};
            deriv(d, xyz[i], xE[i][d], 1, A, nullptr, &m.T[d][i], d == 2 && i == 1);
        }
    // products, :298-310
    const double *xE1[3] = {xE[0][0], xE[0][1], xE[0][2]}, *xE2[3] = {xE[1][0], xE[1][1], xE[1][2]};
    hip_check(launch_metric_products(xE1, xE2, y, z, Va, Vb, Vc, N, st), "metric products");
    // cross derivatives, :348-1001: a = xE1j y, b = xE1j z, c = xE2j z
    for (int k : {1, 0, 2})
        for (int fam = 0; fam < 3; ++fam)
            for (int j = 0; j < 3; ++j) {
                if (j == k) continue;
                double *const *V = fam == 0 ? Va : fam == 1 ? Vb : Vc;
                double *const(*dV)[3] = fam == 0 ? dVa : fam == 1 ? dVb : dVc;
                const int comp = fam == 0 ? 1 : 2;
                const double *A[1] = {xyz[comp]};
                const double *B[1] = {fam == 2 ? xE2[j] : xE1[j]};
                deriv(k, V[j], dV[j][k], 1, A, B, &m.T[k][comp], k == 0 && fam == 2 && j == 2);
            }
    CofactorArgs ca{};
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) { ca.dVa[j][k] = dVa[j][k]; ca.dVb[j][k] = dVb[j][k]; ca.dVc[j][k] = dVc[j][k]; }
    ca.x = x; ca.y = y; ca.z = z;
    double *Jm[9];
    for (int k = 0; k < 9; ++k) Jm[k] = out[1 + k];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) ca.Jm[i][j] = Jm[3 * i + j];
        ca.pre[i] = pre[i];
    }
    hip_check(launch_metric_cofactors(ca, N, st), "metric cofactors");
    // Jacobian determinant, :1057-1298
    for (int k : {1, 0, 2}) {
        const double *B[3] = {Jm[3 * k], Jm[3 * k + 1], Jm[3 * k + 2]};
        deriv(k, pre[k], Jdet[k], 3, xyz, B, m.T[k], false);
    }
    const double *jd[3] = {Jdet[0], Jdet[1], Jdet[2]};
    hip_check(launch_metric_final(jd, out[0], Jm, N, st), "metric determinant");
    hip_check(hipStreamSynchronize(st), "metrics");
}

}  // namespace mc3d
