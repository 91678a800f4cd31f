// Host orchestration of the solver step (see solver.hpp).
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "../dsm.h"
#include "../stats.h"
#include "closure.hpp"
#include "mesh.hpp"
#include "dump.hpp"
#include "../lineop_api.h"
#include "../slice.h"

namespace mc3d {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Error(MC3D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

static void mc3d_check(int rc, const char *what) {
    if (rc != MC3D_OK) throw Error(rc, std::string(what) + ": " + mc3d_last_error());
}

Domain::Domain(const Options &o) {
    gNx = o.Nx; gNy = o.Ny; gNz = o.Nz;
    gN = (long)gNx * gNy * gNz;
    // x[i] = i*L/N (periodic) or i/(N-1)*L, and dx = x[1]-x[0] with L = 1
    auto spacing = [](int n, Options::BCType t) {
        double x0 = 0.0, x1 = (t == Options::PERIODIC_SOLVE) ? 1.0 * 1.0 / (double)n
                                                             : ((1.0) / ((double)n - 1.0)) * 1.0;
        return x1 - x0;
    };
    dx = spacing(gNx, o.bcXType);
    dy = spacing(gNy, o.bcYType);
    dz = spacing(gNz, o.bcZType);
}

// ---------------------------------------------------------------------------
LineDerivative::LineDerivative(int scheme, const Domain &d, Options::BCType bc, Direct dir,
                               hipStream_t st, const LineTune &shape)
    : scheme_(scheme), st_(st) {
    Nx = d.gNx; Ny = d.gNy; Nz = d.gNz;
    currentDir = dir;
    bcType = bc;
    N = dir == DIRX ? Nx : dir == DIRY ? Ny : Nz;
    dd = dir == DIRX ? d.dx : dir == DIRY ? d.dy : d.dz;
    lhsBandwidth = scheme == MC3D_PENTA10 ? BW5 : scheme == MC3D_PADE6 ? BW3 : BW3;
    rhsBandwidth = scheme == MC3D_PENTA10 ? BW7 : scheme == MC3D_PADE6 ? BW5 : BW3;
    const int dims[3] = {Nx, Ny, Nz};
    const int b = bc == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET;
    const int bc3[3] = {b, b, b};
    mc3d_check(lineop_create(kDeriv, scheme, (int)dir, bc3, dims, dd, 0.0, d.members, shape, &op_),
               "derivative setup");
}

LineDerivative::~LineDerivative() { mc3d_lineop_destroy(op_); }

void LineDerivative::calc1stDerivField(const double *in, double *out) {
    calc1stDerivFields(1, &in, &out, false);
}

void LineDerivative::calc1stDerivField_TPB(const double *in, double *out, const double *Nm2,
                                           const double *Nm1, const double *Np1, const double *Np2) {
    const double *w[6] = {nullptr, Nm2, Nm1, Np1, Np2, nullptr};
    mc3d_check(mc3d_deriv_apply_tpb(op_, in, out, w, st_), "transformed-periodic derivative");
}

void LineDerivative::calc1stDerivField_TPB(const double *in, double *out, const double *Nm3,
                                           const double *Nm2, const double *Nm1, const double *Np1,
                                           const double *Np2, const double *Np3) {
    const double *w[6] = {Nm3, Nm2, Nm1, Np1, Np2, Np3};
    mc3d_check(mc3d_deriv_apply_tpb(op_, in, out, w, st_), "transformed-periodic derivative");
}

// src/Pade6.cpp:395-397, src/Penta10.cpp:483-485, src/CD2.cpp:183-185
double LineDerivative::calcNeumann(const double *f) {
    if (scheme_ == MC3D_CD2) return (f[0] * 4.0 - f[1] * 1.0) / 3.0;
    return (f[0] * 360.0 - f[1] * 450.0 + f[2] * 400.0 - f[3] * 225.0 + f[4] * 72.0 - f[5] * 10.0) / 147.0;
}

void LineDerivative::calc1stDerivFields(int n, const double *const *in, double *const *out,
                                        bool accumulate) {
    int t = -1;
    if (timer && timer->enabled())
        t = timer->begin(lineop_kernel_name(op_), (double)n * (accumulate ? 24.0 : 16.0) * Nx * Ny * Nz);
    mc3d_check(lineop_launch_accumulate(op_, in, out, n, accumulate ? 1 : 0, st_), "derivative");
    if (timer) timer->end(t);
}

LineFilter::LineFilter(int kind, double alphaF, const Domain &d, const int bc3[3],
                       AbstractDerivatives::Direct dir, hipStream_t st, const LineTune &shape)
    : st_(st) {
    Nx = d.gNx; Ny = d.gNy; Nz = d.gNz;
    currentDir = dir;
    bcType = bc3[dir] == MC3D_PERIODIC ? Options::PERIODIC_SOLVE : Options::DIRICHLET_SOLVE;
    N = dir == AbstractDerivatives::DIRX ? Nx : dir == AbstractDerivatives::DIRY ? Ny : Nz;
    const int dims[3] = {Nx, Ny, Nz};
    mc3d_check(lineop_create(kFilter, kind, (int)dir, bc3, dims, 1.0, alphaF, d.members, shape, &op_),
               "filter setup");
}

LineFilter::~LineFilter() { mc3d_lineop_destroy(op_); }

void LineFilter::filterField(const double *in, double *out) { filterFields(1, &in, &out); }

void LineFilter::filterFields(int n, const double *const *in, double *const *out) {
    int t = -1;
    if (timer && timer->enabled()) t = timer->begin(lineop_kernel_name(op_), (double)n * 16.0 * Nx * Ny * Nz);
    mc3d_check(lineop_launch_accumulate(op_, in, out, n, 0, st_), "filter");
    if (timer) timer->end(t);
}

// ---------------------------------------------------------------------------
// src/AbstractSGS.cpp:13-164
void AbstractSGS::doAveraging(const double *phi, double *phiF) {
    const Domain &d = cs->dom;
    switch (musgsAvgType) {
        case Options::LOCAL: {
            double *t = cs->tmp[9];
            filtY->filterField(phi, t);
            filtZ->filterField(t, t);
            filtX->filterField(t, phiF);
            break;
        }
        case Options::GLOBAL:
            hip_check(launch_global_mean(phi, phiF, cs->cfl_partial, cs->N, cs->st), "global mean");
            break;
        case Options::XI1_AVG:
        case Options::XI2_AVG:
        case Options::XI3_AVG: {
            const int dir = cs->axis_of[musgsAvgType == Options::XI1_AVG ? 0 : musgsAvgType == Options::XI2_AVG ? 1 : 2];
            hip_check(launch_xi_average(phi, phiF, dir, d.gNx, d.gNy, d.gNz, cs->st), "line average");
            break;
        }
        default:
            if (phi != phiF)
                hip_check(hipMemcpyAsync(phiF, phi, sizeof(double) * cs->N, hipMemcpyDeviceToDevice, cs->st),
                          "copy");
            break;
    }
}

// src/VremanSGS.hpp:15-41
VremanSGS::VremanSGS(CurvilinearCSolver *c) {
    cs = c;
    musgsAvgType = c->opt.musgsAvgType;
    const int bc3[3] = {c->opt.bcXType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        c->opt.bcYType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        c->opt.bcZType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET};
    auto axis = [c](int d) { return (AbstractDerivatives::Direct)c->axis_of[d]; };
    filtX.reset(new Compact10Filter(-0.49, c->dom, bc3, axis(0), c->st));
    filtY.reset(new Compact10Filter(-0.49, c->dom, bc3, axis(1), c->st));
    filtZ.reset(new Compact10Filter(-0.49, c->dom, bc3, axis(2), c->st));
    for (auto *f : {filtX.get(), filtY.get(), filtZ.get()}) static_cast<LineFilter *>(f)->timer = &c->timer;
    mu_sgs = c->alloc_field();
}

void VremanSGS::getSGSViscosity(const double *const dU[3], const double *const dV[3],
                                const double *const dW[3], const double *rho) {
    cs->timed("k_vreman", 21.0, [&] {
        hip_check(launch_vreman(dU, dV, dW, cs->msh, rho, mu_sgs, cs->dom.dx, cs->dom.dy, cs->dom.dz,
                                cs->N, cs->st),
                  "vreman");
    });
    doAveraging(mu_sgs, mu_sgs);
}

// ---------------------------------------------------------------------------
// src/DSMSGS.hpp:34-111
DSMSGS::DSMSGS(CurvilinearCSolver *c, bool fallback_only) {
    cs = c;
    isDSM = !fallback_only;
    musgsAvgType = c->opt.musgsAvgType;
    useTaukk = fallback_only ? false : c->opt.useTaukk;
    coeffRangeFlag = fallback_only ? false : c->opt.dumpCoeffRange;
    const int bc3[3] = {c->opt.bcXType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        c->opt.bcYType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        c->opt.bcZType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET};
    std::unique_ptr<AbstractFilter> *fv[3] = {&filtX, &filtY, &filtZ};
    for (int d = 0; d < 3; ++d) {
        auto dir = (AbstractDerivatives::Direct)c->axis_of[d];
        // src/DSMSGS.hpp:57-67: 1 = Pade-type, 2 = explicit
        if (c->opt.testFilterType == 1) fv[d]->reset(new CommN5PadeTestFilter(c->dom, bc3, dir, c->st));
        else fv[d]->reset(new CommN5ExpTestFilter(c->dom, bc3, dir, c->st));
        static_cast<LineFilter *>(fv[d]->get())->timer = &c->timer;
    }
    mu_sgs = c->alloc_field();
    taukk = c->alloc_field();
    k_sgs = c->alloc_field();
    if (coeffRangeFlag) { C = c->alloc_field(); CI = c->alloc_field(); Prt = c->alloc_field(); }
    for (auto &f : S) f = c->alloc_field();
    Smag = c->alloc_field(); rho_hat = c->alloc_field(); Smag_hat = c->alloc_field();
    rhoU_hat = c->alloc_field(); rhoV_hat = c->alloc_field(); rhoW_hat = c->alloc_field();
    for (int k = 1; k <= 18; ++k) work[k] = c->alloc_field();
    for (auto &f : gT) f = c->alloc_field();
    for (auto &f : ftmp_) f = c->alloc_field();
}

void DSMSGS::filterQuantity(int n, const double *const *in, double *const *out) {
    for (int b = 0; b < n; b += 16) {
        const int m = n - b < 16 ? n - b : 16;
        const double *tc[16];
        for (int k = 0; k < m; ++k) tc[k] = ftmp_[k];
        filtY->filterFields(m, in + b, ftmp_);
        filtZ->filterFields(m, tc, ftmp_);
        filtX->filterFields(m, tc, out + b);
    }
}

void DSMSGS::average(int n, const double *const *in, double *const *out) {
    if (musgsAvgType == Options::LOCAL) filterQuantity(n, in, out);
    else
        for (int k = 0; k < n; ++k) doAveraging(in[k], out[k]);
}

// src/DSMSGS.hpp:124-343
void DSMSGS::calcMuSGSTaukk(const double *const dU[3], const double *const dV[3],
                            const double *const dW[3], const double *rho, const double *rhoU,
                            const double *rhoV, const double *rhoW) {
    const long N = cs->N;
    hipStream_t st = cs->st;
    double **w = work;
    DsmStrainArgs sa{};
    for (int d = 0; d < 3; ++d) { sa.dU[d] = dU[d]; sa.dV[d] = dV[d]; sa.dW[d] = dW[d]; }
    sa.m = cs->msh;
    sa.rho = rho;
    for (int k = 0; k < 6; ++k) { sa.S[k] = S[k]; sa.M2[k] = w[1 + k]; }
    sa.Smag = Smag;
    cs->timed("k_dsm_strain", 32.0, [&] { hip_check(launch_dsm_strain(sa, N, st), "dsm strain"); });
    {   // M_ij (work7..12), rho_hat, S_ij hat (work1..6), Smag_hat
        const double *in[14] = {w[1], w[2], w[3], w[4], w[5], w[6], rho, S[0], S[1], S[2], S[3], S[4], S[5], Smag};
        double *out[14] = {w[7], w[8], w[9], w[10], w[11], w[12], rho_hat, w[1], w[2], w[3], w[4], w[5], w[6], Smag_hat};
        filterQuantity(14, in, out);
    }
    DsmMLArgs ma{};
    for (int k = 0; k < 6; ++k) { ma.M[k] = w[7 + k]; ma.S_hat[k] = w[1 + k]; ma.Lt[k] = w[13 + k]; }
    ma.rho_hat = rho_hat; ma.Smag_hat = Smag_hat;
    ma.rho = rho; ma.rhoU = rhoU; ma.rhoV = rhoV; ma.rhoW = rhoW;
    cs->timed("k_dsm_ml", 30.0, [&] { hip_check(launch_dsm_ml(ma, N, st), "dsm M/L"); });
    {   // L_ij (work1..6), momentum hats
        const double *in[9] = {w[13], w[14], w[15], w[16], w[17], w[18], rhoU, rhoV, rhoW};
        double *out[9] = {w[1], w[2], w[3], w[4], w[5], w[6], rhoU_hat, rhoV_hat, rhoW_hat};
        filterQuantity(9, in, out);
    }
    DsmLArgs la{};
    for (int k = 0; k < 6; ++k) la.L[k] = w[1 + k];
    la.rhoU_hat = rhoU_hat; la.rhoV_hat = rhoV_hat; la.rhoW_hat = rhoW_hat; la.rho_hat = rho_hat;
    la.rho = rho; la.Smag = Smag;
    la.CIdenom = useTaukk ? w[13] : nullptr;
    cs->timed("k_dsm_l", useTaukk ? 19.0 : 16.0, [&] { hip_check(launch_dsm_l(la, N, st), "dsm L"); });
    if (useTaukk) {
        const double *in[1] = {w[13]};
        double *out[1] = {w[14]};
        filterQuantity(1, in, out);  // alpha_hat
    }
    DsmContractArgs ca{};
    for (int k = 0; k < 6; ++k) { ca.L[k] = w[1 + k]; ca.M[k] = w[7 + k]; }
    ca.rho_hat = rho_hat; ca.Smag_hat = Smag_hat;
    ca.alpha_hat = useTaukk ? w[14] : nullptr;
    ca.CIdenom = w[13];
    ca.Lkk = w[15]; ca.LijMij = w[17]; ca.MijMij = w[18];
    cs->timed("k_dsm_contract", useTaukk ? 19.0 : 15.0,
              [&] { hip_check(launch_dsm_contract(ca, N, st), "dsm contractions"); });
    if (useTaukk) {
        const double *in[4] = {w[15], w[13], w[17], w[18]};
        double *out[4] = {w[7], w[8], w[9], w[10]};
        average(4, in, out);
    } else {
        const double *in[2] = {w[17], w[18]};
        double *out[2] = {w[9], w[10]};
        average(2, in, out);
    }
    DsmMuArgs ua{};
    ua.rho = rho; ua.Smag = Smag; ua.LijMij_avg = w[9]; ua.MijMij_avg = w[10];
    ua.Lkk_avg = w[7]; ua.CIdenom_avg = w[8]; ua.J = cs->msh.J;
    ua.mu_sgs = mu_sgs; ua.taukk = taukk; ua.C = C; ua.CI = CI;
    ua.use_taukk = useTaukk ? 1 : 0;
    ua.dx = cs->dom.dx; ua.dy = cs->dom.dy; ua.dz = cs->dom.dz;
    cs->timed("k_dsm_mu", 6.0 + (useTaukk ? 2.0 : 0.0) + (C ? 3.0 : 0.0),
              [&] { hip_check(launch_dsm_mu(ua, N, st), "dsm mu"); });
}

// src/DSMSGS.hpp:345-444
void DSMSGS::calcKSGS(const double *const dT[3], const double *rho, const double *rhoU,
                      const double *rhoV, const double *rhoW, const double *T) {
    const long N = cs->N;
    hipStream_t st = cs->st;
    double **w = work;
    DsmNKArgs na{};
    for (int d = 0; d < 3; ++d) { na.dT[d] = dT[d]; na.N[d] = w[1 + d]; na.K[d] = w[7 + d]; na.gT[d] = gT[d]; }
    na.m = cs->msh;
    na.rho = rho; na.Smag = Smag; na.rhoU = rhoU; na.rhoV = rhoV; na.rhoW = rhoW; na.T = T;
    na.rhoT = w[16];
    cs->timed("k_dsm_nk", 28.0, [&] { hip_check(launch_dsm_nk(na, N, st), "dsm N/K"); });
    {   // NRHS (work4..6), KRHS (work10..12), rhoT_hat (work17), gradT hats (work13..15)
        const double *in[10] = {w[1], w[2], w[3], w[7], w[8], w[9], w[16], gT[0], gT[1], gT[2]};
        double *out[10] = {w[4], w[5], w[6], w[10], w[11], w[12], w[17], w[13], w[14], w[15]};
        filterQuantity(10, in, out);
    }
    DsmNNArgs nn{};
    nn.rho_hat = rho_hat; nn.Smag_hat = Smag_hat;
    for (int d = 0; d < 3; ++d) { nn.g_hat[d] = w[13 + d]; nn.NRHS[d] = w[4 + d]; nn.KRHS[d] = w[10 + d]; }
    nn.rhoU_hat = rhoU_hat; nn.rhoV_hat = rhoV_hat; nn.rhoW_hat = rhoW_hat; nn.rhoT_hat = w[17];
    nn.rho = rho;
    nn.NiNi = w[4]; nn.NiKi = w[5];
    cs->timed("k_dsm_nn", 18.0, [&] { hip_check(launch_dsm_nn(nn, N, st), "dsm NiNi/NiKi"); });
    {
        const double *in[2] = {w[4], w[5]};
        double *out[2] = {w[6], w[10]};
        average(2, in, out);
    }
    DsmKArgs ka{};
    ka.NiNi_avg = w[6]; ka.NiKi_avg = w[10]; ka.rho = rho; ka.Smag = Smag; ka.C = C; ka.J = cs->msh.J;
    ka.k_sgs = k_sgs; ka.Prt = Prt;
    ka.cp = cs->ig.cp; ka.dx = cs->dom.dx; ka.dy = cs->dom.dy; ka.dz = cs->dom.dz;
    cs->timed("k_dsm_ksgs", Prt ? 8.0 : 5.0, [&] { hip_check(launch_dsm_ksgs(ka, N, st), "dsm k_sgs"); });
}

// ---------------------------------------------------------------------------
void AbstractRK::executeSolverLoop() {
    while (!cs->endFlag) step();
}

void AbstractRK::step() {
    advance();
    cs->postStep();
}

// src/TVDRK3.hpp:135-176, src/RK4.hpp:167-218, src/KenRK4.hpp:154-215,
// src/LSLDDRK4.hpp:324-460: preStep, the stages, then (in step) postStep
void AbstractRK::advance() {
    const int S = stages();
    cs->rkLast = false;
    cs->preStep();
    for (int s = 1; s <= S; ++s) cs->run_stage(this, s, S);
    finishStep();
}

void CurvilinearCSolver::run_stage(AbstractRK *r, int s, int S) {
    rkStep = s;
    if (s == S) rkLast = true;
    preSubStep();
    solveEqnSet();
    postSubStep();
    r->updateConservedData();
    updateData();
}

// src/TVDRK3.hpp:28-37 coefficients, :50-133 stages
void TVDRK3::updateConservedData() {
    switch (cs->rkStep) {
        case 1: cs->stage_update(kTvd1, 0.0, 0.0, 1.0); break;
        case 2: cs->stage_update(kTvd2, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0); break;
        default: cs->stage_update(kTvd3, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0); break;
    }
}

// src/RK4.hpp:27-34, :47-165
void RK4::updateConservedData() {
    const double a[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double b[3] = {1.0 / 2.0, 1.0 / 2.0, 1.0};
    const int s = cs->rkStep;
    if (s == 1) cs->stage_update(kRk4First, a[0], b[0], 0.0);
    else if (s < 4) cs->stage_update(kRk4Mid, a[s - 1], b[s - 1], 0.0);
    else cs->stage_update(kRk4Last, a[3], 0.0, 0.0);
}

// src/KenRK4.hpp:30-40, :53-215
static const double kKenA[5] = {0.0, -6234157559845.0 / 12983515589748.0,
                                -6194124222391.0 / 4410992767914.0,
                                -31623096876824.0 / 15682348800105.0,
                                -12251185447671.0 / 11596622555746.0};
static const double kKenB[5] = {494393426753.0 / 4806282396855.0, 4047970641027.0 / 5463924506627.0,
                                9795748752853.0 / 13190207949281.0, 4009051133189.0 / 8539092990294.0,
                                1348533437543.0 / 7166442652324.0};

static void lowstorage_update(CurvilinearCSolver *cs, const double *A, const double *B, int S) {
    const int s = cs->rkStep;
    if (s == 1) cs->stage_update(kLsFirst, 0.0, B[0], 0.0);
    else if (s < S) cs->stage_update(kLsMid, A[s - 1], B[s - 1], 0.0);
    else cs->stage_update(kLsLast, A[s - 1], B[s - 1], 0.0);
}

void KenRK4::updateConservedData() { lowstorage_update(cs, kKenA, kKenB, 5); }

// src/LSLDDRK4.hpp:40-78, :93-460
static const double kLs1A[5] = {0.0, -0.60512264332862261228, -2.04375640234761394333,
                                -0.74069990637544192841, -4.42317651302968168941};
static const double kLs1B[5] = {0.26874543888713438496, 0.8014706973220802933, 0.50515704269422722538,
                                0.56235680379000296407, 0.05900655127758823335};
static const double kLs2A[6] = {0.0, -0.44127377153877382565, -1.073982008079781868,
                                -1.7063570791256758809, -2.7979293162682443056, -4.0913537120919160454};
static const double kLs2B[6] = {0.11584888181285561688, 0.37287699051652864918, 0.73795368921435295698,
                                0.57981109366311039538, 1.031284991300145194, 0.15};

void LSLDDRK4::updateConservedData() {
    if (rkDualStepStep == 1) lowstorage_update(cs, kLs1A, kLs1B, 5);
    else lowstorage_update(cs, kLs2A, kLs2B, 6);
}

// ---------------------------------------------------------------------------
// src/CurvilinearCSolver.hpp:120-245
CurvilinearCSolver::CurvilinearCSolver(const Options &o, hipStream_t s, int members, const double *mu_members)
    : opt(o), dom(o), st(s), timer(s) {
    if (members < 1) throw Error(MC3D_EINVAL, "member count must be positive");
    dom.members = members;
    N = dom.gN * members;
    if (members > 1) {
        const bool periodic = opt.bcXType == Options::PERIODIC_SOLVE && opt.bcYType == Options::PERIODIC_SOLVE &&
                              opt.bcZType == Options::PERIODIC_SOLVE;
        if (!periodic) throw Error(MC3D_ECONFIG, "stacked members must be periodic boxes");
        if (opt.lesModel == Options::DSM || opt.lesModel == Options::LEARNED)
            throw Error(MC3D_ECONFIG, "stacked members take no LES model or VREMAN");
        if (opt.lesModel == Options::VREMAN && opt.musgsAvgType != Options::NONE &&
            opt.musgsAvgType != Options::LOCAL)
            throw Error(MC3D_ECONFIG, "stacked members take LESAVERAGING NONE or LOCAL");
        if (N > 268435455L) throw Error(MC3D_EINVAL, "stacked members above 2^28 points per field");
        if (opt.velocityStats || opt.thermoStats)
            throw Error(MC3D_ECONFIG, "stacked members keep no running statistics");
    }
    timeStep = opt.timeStep;
    initGas();
    // field slab
    const int nfields = 10 + 20 + 10 + 12 + 15 + 5 + 10 + 8;
    slab_bytes_ = sizeof(double) * (size_t)N * nfields + 4096;
    hip_check(hipMalloc(&slab_, slab_bytes_), "field allocation");
    for (int k = 0; k < 10; ++k) Jm[k] = alloc_field();
    msh = Metrics{Jm[0], Jm[1], Jm[2], Jm[3], Jm[4], Jm[5], Jm[6], Jm[7], Jm[8], Jm[9]};
    for (int q = 0; q < 5; ++q) { r1.q[q] = alloc_field(); rk_.q[q] = alloc_field(); k2.q[q] = alloc_field(); r2.q[q] = alloc_field(); }
    pr = Prim{alloc_field(), alloc_field(), alloc_field(), alloc_field(), alloc_field(),
              alloc_field(), alloc_field(), alloc_field(), alloc_field(), alloc_field()};
    for (int d = 0; d < 3; ++d) { dU[d] = alloc_field(); dV[d] = alloc_field(); dW[d] = alloc_field(); dT[d] = alloc_field(); }
    for (int e = 0; e < 15; ++e) pre[e] = alloc_field();
    for (int e = 0; e < 5; ++e) div[e] = alloc_field();
    for (int e = 0; e < 10; ++e) tmp[e] = alloc_field();
    // zero everything: k2 registers must start at zero (momentum increments accumulate)
    hip_check(hipMemsetAsync(slab_, 0, slab_bytes_, st), "field clear");
    void *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(double) * std::max<size_t>(4096, (size_t)cfl_blocks(dom.gN) * members)),
              "scratch allocation");
    extra_.push_back(p);
    cfl_partial = static_cast<double *>(p);
    hip_check(hipMalloc(&p, sizeof(StepScalars) * members), "scalar allocation");
    extra_.push_back(p);
    sc = static_cast<StepScalars *>(p);
    const std::vector<StepScalars> s0((size_t)members, StepScalars{opt.dt, opt.CFL, opt.time, 0.0});
    hip_check(hipMemcpy(sc, s0.data(), sizeof(StepScalars) * members, hipMemcpyHostToDevice), "scalar init");
    if (members > 1) {
        // per-member MU_REF (the options' value for every member unless given)
        std::vector<double> mu((size_t)members, opt.mu_ref);
        if (mu_members) mu.assign(mu_members, mu_members + members);
        hip_check(hipMalloc(&p, sizeof(double) * members), "member allocation");
        extra_.push_back(p);
        hip_check(hipMemcpy(p, mu.data(), sizeof(double) * members, hipMemcpyHostToDevice), "member upload");
        ig.mu_member = static_cast<const double *>(p);
        ig.member_pts = (unsigned)dom.gN;
        mu_members_ = mu;
    }

    setupBoundaries();
    // src/CurvilinearCSolver.hpp:162-217
    const Options::BCType bt[3] = {opt.bcXType, opt.bcYType, opt.bcZType};
    const Options::FDType ft[3] = {opt.xFDType, opt.yFDType, opt.zFDType};
    std::unique_ptr<AbstractDerivatives> *dv[3] = {&derivX, &derivY, &derivZ};
    for (int d = 0; d < 3; ++d) {
        auto dir = (AbstractDerivatives::Direct)d;
        if (ft[d] == Options::CD2) dv[d]->reset(new CD2(dom, bt[d], dir, st));
        else if (ft[d] == Options::PADE6) dv[d]->reset(new Pade6(dom, bt[d], dir, st));
        else dv[d]->reset(new Penta10(dom, bt[d], dir, st));
    }
    const int bc3[3] = {bt[0] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        bt[1] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        bt[2] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET};
    std::unique_ptr<AbstractFilter> *fv[3] = {&filtX, &filtY, &filtZ};
    for (int d = 0; d < 3; ++d) {
        auto dir = (AbstractDerivatives::Direct)d;
        if (opt.filterType == Options::COMPACT8) fv[d]->reset(new Compact8Filter(opt.alphaF, dom, bc3, dir, st));
        else fv[d]->reset(new Compact10Filter(opt.alphaF, dom, bc3, dir, st));
    }
    for (auto *d : {derivX.get(), derivY.get(), derivZ.get()}) static_cast<LineDerivative *>(d)->timer = &timer;
    for (auto *f : {filtX.get(), filtY.get(), filtZ.get()}) static_cast<LineFilter *>(f)->timer = &timer;
    // src/CurvilinearCSolver.hpp:232-242
    if (opt.lesModel == Options::VREMAN) les.reset(new VremanSGS(this));
    else if (opt.lesModel == Options::DSM) les.reset(new DSMSGS(this));
    else if (opt.lesModel == Options::LEARNED) les.reset(new LearnedSGS(this));
    if (opt.velocityStats || opt.thermoStats) stats.reset(new Stats(this));
    // src/MPI_3DCompact.cpp:164-176
    switch (opt.rkType) {
        case Options::TVDRK3: rk.reset(new TVDRK3(this)); break;
        case Options::RK4: rk.reset(new RK4(this)); break;
        case Options::KENRK4: rk.reset(new KenRK4(this)); break;
        default: rk.reset(new LSLDDRK4(this)); break;
    }
}

// src/IdealGas.hpp:35-48
void CurvilinearCSolver::initGas() {
    ig.gamma = 1.4;
    ig.Pr = 0.7;
    const double p_ref = 1.0 / ig.gamma, rho_ref = 1.0;
    ig.T_ref = 1.0;
    ig.mu_ref = opt.mu_ref;
    ig.R_gas = p_ref / rho_ref / ig.T_ref;
    ig.cp = ig.R_gas * ig.gamma / (ig.gamma - 1.0);
    ig.mu_member = nullptr;
    ig.member_pts = 0;
}

namespace {
// the options with the grid and boundary types in the axis order of a layout
// (MC3D_XPENCIL: x, y, z; YPENCIL: y, z, x; ZPENCIL: z, x, y)
Options rotated(const Options &o, int r) {
    if (r < MC3D_XPENCIL || r > MC3D_ZPENCIL) throw Error(MC3D_EINVAL, "unknown layout");
    Options v = o;
    const int n[3] = {o.Nx, o.Ny, o.Nz};
    const Options::BCType b[3] = {o.bcXType, o.bcYType, o.bcZType};
    v.Nx = n[r % 3]; v.Ny = n[(1 + r) % 3]; v.Nz = n[(2 + r) % 3];
    v.bcXType = b[r % 3]; v.bcYType = b[(1 + r) % 3]; v.bcZType = b[(2 + r) % 3];
    return v;
}
}  // namespace

CurvilinearCSolver::CurvilinearCSolver(const Options &o, hipStream_t s, SgsOnly t)
    : opt(rotated(o, t.layout)), dom(opt), st(s), timer(s) {
    if (opt.lesModel != Options::VREMAN && opt.lesModel != Options::DSM && opt.lesModel != Options::LEARNED)
        throw Error(MC3D_ECONFIG, "an SGS handle needs LESMODEL VREMAN, DSM or LEARNED");
    if (!t.J) throw Error(MC3D_EINVAL, "null Jacobian field");
    N = dom.gN;
    timeStep = opt.timeStep;
    // the filter width (dxi deta dzeta / J)^(1/3) multiplies the spacings in x, y, z order
    const Domain phys(o);
    dom.dx = phys.dx; dom.dy = phys.dy; dom.dz = phys.dz;
    for (int d = 0; d < 3; ++d) axis_of[d] = (d - t.layout + 3) % 3;
    initGas();
    slab_bytes_ = sizeof(double) * (size_t)N + 4096;
    hip_check(hipMalloc(&slab_, slab_bytes_), "field allocation");
    tmp[9] = alloc_field();
    msh = Metrics{t.J, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(double) * std::max<size_t>(4096, (size_t)cfl_blocks(dom.gN))),
              "scratch allocation");
    extra_.push_back(p);
    cfl_partial = static_cast<double *>(p);
    if (opt.lesModel == Options::VREMAN) les.reset(new VremanSGS(this));
    else if (opt.lesModel == Options::DSM) les.reset(new DSMSGS(this));
    else les.reset(new LearnedSGS(this));
}

CurvilinearCSolver::~CurvilinearCSolver() {
    les.reset();
    (void)hipStreamSynchronize(st);
    for (void *p : extra_) (void)hipFree(p);
    if (slab_) (void)hipFree(slab_);
    if (slice_buf_) (void)hipFree(slice_buf_);
}

double *CurvilinearCSolver::alloc_field() {
    const size_t bytes = sizeof(double) * (size_t)N;
    const size_t off = (slab_used_ + 255) & ~size_t(255);
    if (off + bytes > slab_bytes_) {
        void *p = nullptr;
        hip_check(hipMalloc(&p, bytes), "field allocation");
        hip_check(hipMemsetAsync(p, 0, bytes, st), "field clear");
        extra_.push_back(p);
        return static_cast<double *>(p);
    }
    slab_used_ = off + bytes;
    return reinterpret_cast<double *>(static_cast<char *>(slab_) + off);
}

void CurvilinearCSolver::set_metrics_host(const double *const m[10]) {
    for (int k = 0; k < 10; ++k)
        hip_check(hipMemcpyAsync(Jm[k], m[k], sizeof(double) * N, hipMemcpyHostToDevice, st), "metric upload");
    hip_check(hipStreamSynchronize(st), "metric upload");
}

// src/CurvilinearCSolver_BC.cpp:3-63, then the wall initial fixes (:63-322) and
// the sponge averages (:324-345)
void CurvilinearCSolver::setInitialConditions_host(const double *rho0, const double *U0,
                                                   const double *V0, const double *W0,
                                                   const double *p0) {
    const double *src[5] = {rho0, U0, V0, W0, p0};
    double *d[5] = {tmp[0], tmp[1], tmp[2], tmp[3], tmp[4]};
    for (int k = 0; k < 5; ++k)
        hip_check(hipMemcpyAsync(d[k], src[k], sizeof(double) * N, hipMemcpyHostToDevice, st), "IC upload");
    hip_check(launch_init_state(d[0], d[1], d[2], d[3], d[4], r1, pr, msh, ig, N, st), "initial state");
    wallInitialConditions();
    initSpongeAverages();
    hip_check(hipStreamSynchronize(st), "initial state");
}

// member b of the stacked boxes: the initial conversion of its slice with its
// own MU_REF (the same kernel as the whole-field form)
void CurvilinearCSolver::setMemberInitial(int b, const double *rho0, const double *U0, const double *V0,
                                          const double *W0, const double *p0) {
    if (b < 0 || b >= dom.members) throw Error(MC3D_EINVAL, "member index out of range");
    if (dom.members == 1) return setInitialConditions_host(rho0, U0, V0, W0, p0);
    const long off = (long)b * dom.gN;
    const double *src[5] = {rho0, U0, V0, W0, p0};
    for (int k = 0; k < 5; ++k)
        hip_check(hipMemcpyAsync(tmp[k] + off, src[k], sizeof(double) * dom.gN, hipMemcpyHostToDevice, st),
                  "IC upload");
    Cons r{};
    for (int q = 0; q < 5; ++q) r.q[q] = r1.q[q] + off;
    Prim p{pr.U + off, pr.V + off, pr.W + off, pr.p + off, pr.T + off,
           pr.mu + off, pr.sos + off, pr.Ucurv + off, pr.Vcurv + off, pr.Wcurv + off};
    Metrics m{msh.J + off, msh.J11 + off, msh.J12 + off, msh.J13 + off, msh.J21 + off,
              msh.J22 + off, msh.J23 + off, msh.J31 + off, msh.J32 + off, msh.J33 + off};
    Gas g = ig;
    g.mu_member = nullptr;
    g.member_pts = 0;
    g.mu_ref = mu_members_[(size_t)b];
    hip_check(launch_init_state(tmp[0] + off, tmp[1] + off, tmp[2] + off, tmp[3] + off, tmp[4] + off, r, p, m, g,
                                dom.gN, st),
              "initial state");
    hip_check(hipStreamSynchronize(st), "initial state");
}

void CurvilinearCSolver::setMemberMetrics(int b, const double *const m[10]) {
    if (b < 0 || b >= dom.members) throw Error(MC3D_EINVAL, "member index out of range");
    const long off = (long)b * dom.gN;
    for (int k = 0; k < 10; ++k)
        hip_check(hipMemcpyAsync(Jm[k] + off, m[k], sizeof(double) * dom.gN, hipMemcpyHostToDevice, st),
                  "metric upload");
    hip_check(hipStreamSynchronize(st), "metric upload");
}

double *CurvilinearCSolver::memberField(int b, const std::string &name) {
    if (b < 0 || b >= dom.members) throw Error(MC3D_EINVAL, "member index out of range");
    double *p = field_by_name(name);
    return p ? p + (long)b * dom.gN : nullptr;
}

StepScalars CurvilinearCSolver::memberScalars(int b) {
    if (b < 0 || b >= dom.members) throw Error(MC3D_EINVAL, "member index out of range");
    StepScalars h{};
    hip_check(hipMemcpyAsync(&h, sc + b, sizeof(h), hipMemcpyDeviceToHost, st), "scalar read");
    hip_check(hipStreamSynchronize(st), "scalar read");
    return h;
}

void CurvilinearCSolver::setLineShapes(const LineTune deriv[3], const LineTune filt[3]) {
    hip_check(hipStreamSynchronize(st), "operator rebuild");
    const Options::BCType bt[3] = {opt.bcXType, opt.bcYType, opt.bcZType};
    const Options::FDType ft[3] = {opt.xFDType, opt.yFDType, opt.zFDType};
    std::unique_ptr<AbstractDerivatives> *dv[3] = {&derivX, &derivY, &derivZ};
    std::unique_ptr<AbstractFilter> *fv[3] = {&filtX, &filtY, &filtZ};
    const int bc3[3] = {bt[0] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        bt[1] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                        bt[2] == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET};
    for (int d = 0; d < 3; ++d) {
        const auto dir = (AbstractDerivatives::Direct)d;
        const int scheme = ft[d] == Options::CD2 ? MC3D_CD2 : ft[d] == Options::PADE6 ? MC3D_PADE6 : MC3D_PENTA10;
        auto *nd = new LineDerivative(scheme, dom, bt[d], dir, st, deriv[d]);
        nd->timer = &timer;
        dv[d]->reset(nd);
        const int kind = opt.filterType == Options::COMPACT8 ? MC3D_COMPACT8 : MC3D_COMPACT10;
        auto *nf = new LineFilter(kind, opt.alphaF, dom, bc3, dir, st, filt[d]);
        nf->timer = &timer;
        fv[d]->reset(nf);
    }
}

// The launches of a step depend on the host state through: the first step's
// time reset, whether this step filters and in which direction order
// (filterTimeStep % 3), and the RK scheme's step form.
int CurvilinearCSolver::stepKey() const {
    if (timeStep == opt.timeStep) return -1;
    const bool filt = (timeStep + 1) % opt.filterStep == 0;
    return rk->variant() * 8 + (filt ? 1 + filterTimeStep % 3 : 0);
}

void CurvilinearCSolver::skipAdvance() {
    const int S = rk->stages();
    timeStep++;
    if (timeStep % opt.filterStep == 0) filterTimeStep += opt.subStepFiltering ? S : 1;
    rkStep = S;
    rkLast = true;
    eos_done_ = false;
    rk->finishStep();
}

// handlePeriodicStuff, src/AbstractSingleBlockMesh.cpp:2019-2085: a periodic
// direction translates by its PERIODICDISP row when the faces are PERIODIC
// (for x the reference reads the y faces, :2034), by zero when they are
// INTERNALLY_PERIODIC; anything else is reported (the reference then leaves
// the translation unset; zero here).
void CurvilinearCSolver::periodicTranslations(double T[3][3]) const {
    const Options::BCType bt[3] = {opt.bcXType, opt.bcYType, opt.bcZType};
    const Options::BCKind lo[3] = {opt.bcX0, opt.bcY0, opt.bcZ0};
    const Options::BCKind hi[3] = {opt.bcX1, opt.bcY1, opt.bcZ1};
    for (int d = 0; d < 3; ++d) {
        for (int e = 0; e < 3; ++e) T[d][e] = 0.0;
        if (bt[d] != Options::PERIODIC_SOLVE) continue;
        const int checked = d == 0 ? 1 : d;
        if (lo[checked] == Options::PERIODIC && hi[checked] == Options::PERIODIC) {
            for (int e = 0; e < 3; ++e) T[d][e] = opt.periodicDisp[d][e];
        } else if (!(lo[d] == Options::INTERNALLY_PERIODIC && hi[d] == Options::INTERNALLY_PERIODIC)) {
            std::printf("  Incompatable periodic bc inputs in %c-direction! \n", "xyz"[d]);
        }
    }
}

// AbstractSingleBlockMesh::solveForJacobians (src/AbstractSingleBlockMesh.cpp:23-1309)
// with this solver's derivative objects and the translations of its options.
void CurvilinearCSolver::solveForJacobians(const double *x, const double *y, const double *z) {
    const Options::BCType bt[3] = {opt.bcXType, opt.bcYType, opt.bcZType};
    const Options::FDType ft[3] = {opt.xFDType, opt.yFDType, opt.zFDType};
    MeshSpec m{};
    m.N[0] = dom.gNx; m.N[1] = dom.gNy; m.N[2] = dom.gNz;
    for (int d = 0; d < 3; ++d) {
        m.periodic[d] = bt[d] == Options::PERIODIC_SOLVE;
        m.bw7[d] = ft[d] == Options::PENTA10;
        if (m.periodic[d] && ft[d] == Options::CD2)
            throw Error(MC3D_ECONFIG, "metrics: CD2 on a periodic direction has no transformed-periodic form "
                                      "(src/AbstractSingleBlockMesh.cpp:94-97)");
    }
    periodicTranslations(m.T);
    AbstractDerivatives *dv[3] = {derivX.get(), derivY.get(), derivZ.get()};
    double *out[10];
    for (int k = 0; k < 10; ++k) out[k] = msh_rw(k);
    solve_for_jacobians(dv, m, x, y, z, out, st);
}

void CurvilinearCSolver::set_mesh_host(const double *x, const double *y, const double *z) {
    void *mem = nullptr;
    hip_check(hipMalloc(&mem, sizeof(double) * 3 * (size_t)N), "mesh upload");
    struct Free {
        void *p;
        ~Free() { (void)hipFree(p); }
    } guard{mem};
    double *d = static_cast<double *>(mem);
    const double *h[3] = {x, y, z};
    for (int k = 0; k < 3; ++k)
        hip_check(hipMemcpyAsync(d + k * N, h[k], sizeof(double) * N, hipMemcpyHostToDevice, st), "mesh upload");
    solveForJacobians(d, d + N, d + 2 * N);
    // the reference reports the metric ranges and identities after forming them
    // (src/AbstractSingleBlockMesh.cpp:1323-1393)
    static const char *names[10] = {"J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33"};
    for (int k = 0; k < 10; ++k) getRange(msh_rw(k), names[k]);
    double range[6];
    metricIdentity(range, true);
    setSpongeFromMesh(x, y, z);
}

// src/MPI_3DCompact.cpp:214-271: rho from the file, U = rhoU / rho (IdealGas
// solveU), p from rhoE (solvep), then setInitialConditions
void CurvilinearCSolver::load_restart_host(const std::string &path) {
    const DumpHeader h = read_dump_header(path);
    if (h.Nx != dom.gNx || h.Ny != dom.gNy || h.Nz != dom.gNz)
        throw Error(MC3D_EINVAL, path + ": grid differs from the solver's");
    std::vector<double> f[5];
    for (auto &v : f) v.resize((size_t)N);
    double *fields[kDumpFields] = {nullptr, nullptr, nullptr, f[0].data(), f[1].data(),
                                   f[2].data(), f[3].data(), f[4].data()};
    read_dump_fields(path, h, fields);
    std::vector<double> U((size_t)N), V((size_t)N), W((size_t)N), p((size_t)N);
    const double g1 = ig.gamma - 1.0;
    for (long i = 0; i < N; ++i) {
        const double rho = f[0][i];
        U[i] = f[1][i] / rho;
        V[i] = f[2][i] / rho;
        W[i] = f[3][i] / rho;
        p[i] = g1 * (f[4][i] - 0.5 * rho * (U[i] * U[i] + V[i] * V[i] + W[i] * W[i]));
    }
    setInitialConditions_host(f[0].data(), U.data(), V.data(), W.data(), p.data());
}

// src/CurvilinearCSolver_Core.cpp:1333-1409
void CurvilinearCSolver::dumpSolution(const std::string &path, const double *const mesh[3]) {
    StepScalars h;
    hip_check(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st), "scalar read");
    std::vector<double> f[5];
    for (int q = 0; q < 5; ++q) {
        f[q].resize((size_t)N);
        hip_check(hipMemcpyAsync(f[q].data(), r1.q[q], sizeof(double) * N, hipMemcpyDeviceToHost, st),
                  "dump read-back");
    }
    hip_check(hipStreamSynchronize(st), "dump read-back");
    DumpHeader d;
    d.timeStep = timeStep;
    d.time = h.time;
    d.Nx = dom.gNx; d.Ny = dom.gNy; d.Nz = dom.gNz;
    const double *fields[kDumpFields] = {mesh[0], mesh[1], mesh[2], f[0].data(), f[1].data(),
                                         f[2].data(), f[3].data(), f[4].data()};
    std::string name = path;
    if (name.empty()) name = (out_dir_.empty() ? std::string() : out_dir_ + "/") + dump_name(timeStep);
    write_dump(name, d, fields);
    // the sponge averages and the statistics go beside it (src/CurvilinearCSolver_Core.cpp:1413-1419)
    const size_t slash = name.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : name.substr(0, slash);
    if (sponge_) dumpSpongeAvg(dir);
    if (stats) stats->dumpStatsFields(dir, timeStep);
}

void CurvilinearCSolver::dumpSpongeAvg(const std::string &dir) {
    const std::string name = (dir.empty() ? std::string() : dir + "/") + "SpongeAvgDump." + std::to_string(timeStep);
    std::ofstream out(name, std::ios::binary | std::ios::trunc);
    if (!out) throw Error(MC3D_EINVAL, "cannot write " + name);
    std::vector<double> buf((size_t)N);
    for (int q = 0; q < 5; ++q) {
        hip_check(hipMemcpyAsync(buf.data(), spgAvg_[q], sizeof(double) * N, hipMemcpyDeviceToHost, st),
                  "sponge read-back");
        hip_check(hipStreamSynchronize(st), "sponge read-back");
        out.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(sizeof(double) * N));
    }
    if (!out) throw Error(MC3D_EINVAL, "write failed: " + name);
}

void CurvilinearCSolver::set_output(const std::string &dir, const double *const mesh[3]) {
    out_dir_ = dir;
    for (int k = 0; k < 3; ++k) {
        if (mesh[k]) mesh_[k].assign(mesh[k], mesh[k] + N);
        else mesh_[k].clear();
    }
    dumps_ = true;
}

void CurvilinearCSolver::addImageOutput(const ImageOutput &img) {
    if (img.width < 1 || img.height < 1) throw Error(MC3D_EINVAL, "image size must be positive");
    if (img.plane < 0 || img.plane > 2) throw Error(MC3D_EINVAL, "image plane must be 0, 1 or 2");
    if (img.interval < 0) throw Error(MC3D_EINVAL, "negative image interval");
    if (img.colormap < GREYSCALE || img.colormap > BWR) throw Error(MC3D_EINVAL, "unknown colormap");
    if (!(img.fraction >= 0.0 && img.fraction <= 1.0)) throw Error(MC3D_EINVAL, "plane fraction must be in [0, 1]");
    if (img.var_name.empty()) throw Error(MC3D_EINVAL, "empty image name");
    if (!field_by_name(img.field)) throw Error(MC3D_EINVAL, "unknown field " + img.field);
    if (dom.members > 1) throw Error(MC3D_ECONFIG, "image outputs take one member");
    images_.push_back(img);
}

void CurvilinearCSolver::slice(const std::string &name, int dir, int index, double *host_out) {
    const double *f = field_by_name(name);
    if (!f) throw Error(MC3D_EINVAL, "unknown field " + name);
    if (dir < 0 || dir > 2) throw Error(MC3D_EINVAL, "slice direction must be 0, 1 or 2");
    const int n[3] = {dom.gNx, dom.gNy, dom.gNz};
    if (index < 0 || index >= n[dir]) throw Error(MC3D_EINVAL, "slice index out of range");
    const size_t np = (size_t)N / n[dir];
    if (np > slice_cap_) {
        if (slice_buf_) hip_check(hipFree(slice_buf_), "slice buffer");
        slice_buf_ = nullptr;
        slice_cap_ = 0;
        hip_check(hipMalloc(&slice_buf_, sizeof(double) * np), "slice buffer");
        slice_cap_ = np;
    }
    hip_check(launch_slice(f, slice_buf_, dir, index, n[0], n[1], n[2], st), "slice");
    hip_check(hipMemcpyAsync(host_out, slice_buf_, sizeof(double) * np, hipMemcpyDeviceToHost, st),
              "slice download");
    hip_check(hipStreamSynchronize(st), "slice download");
}

// src/CurvilinearCSolver_Core.cpp:1442-1492: at the first step every image
// whose interval is 0 or divides the step, later the ones whose interval does
void CurvilinearCSolver::writeImages() {
    if (images_.empty() || images_step_ == timeStep) return;
    const bool first = timeStep == opt.timeStep;
    char step[32];
    std::snprintf(step, sizeof(step), "%06d", timeStep);
    std::vector<double> plane;
    for (const ImageOutput &img : images_) {
        const bool due = img.interval > 0 ? timeStep % img.interval == 0 : first;
        if (!due) continue;
        const int n[3] = {dom.gNx, dom.gNy, dom.gNz};
        const int a1 = img.plane == 0 ? n[1] : img.plane == 1 ? n[2] : n[0];
        const int a2 = img.plane == 0 ? n[2] : img.plane == 1 ? n[0] : n[1];
        plane.resize((size_t)a1 * a2);
        slice(img.field, img.plane, plane_index(img.fraction, n[img.plane]), plane.data());
        const std::vector<unsigned char> rgb = render_plane(plane.data(), a1, a2, img);
        const std::string path = (out_dir_.empty() ? std::string() : out_dir_ + "/") + img.var_name + "." + step + ".png";
        if (!write_png(path, rgb.data(), img.width, img.height))
            std::printf("Warning: could not open png file: %s\n", path.c_str());
    }
    images_step_ = timeStep;
}

// calcDtFromCFL, src/CurvilinearCSolver_Core.cpp:189-237 (images first at the
// first step, src/CurvilinearCSolver.hpp:310-316)
void CurvilinearCSolver::preStep() {
    if (timeStep == opt.timeStep) writeImages();
    timed("k_cfl", 7.0, [&] {
        hip_check(launch_cfl(pr, msh, dom.dx, dom.dy, dom.dz, cfl_partial, dom.gN, dom.members, st), "cfl");
    });
    const bool first = (timeStep == opt.timeStep);
    hip_check(launch_cfl_finish(cfl_partial, cfl_blocks(dom.gN), sc, dom.members, opt.TSType == Options::CONST_CFL,
                                opt.CFL, first ? 1 : 0, opt.time, st),
              "cfl finish");
    timeStep++;
}

// src/CurvilinearCSolver_Core.cpp:468-787
void CurvilinearCSolver::evalSGS() {
    if (les && les->isDSM)
        throw Error(MC3D_ECONFIG, "the DSM changes the pressure through tau_kk; its fields exist only inside a step");
    const int saved = rkStep;
    rkStep = 1;  // the current solution is the *1 register set
    try {
        gradientsAndSGS();
    } catch (...) {
        rkStep = saved;
        throw;
    }
    rkStep = saved;
    hip_check(hipStreamSynchronize(st), "SGS evaluation");
}

void CurvilinearCSolver::sampleClosure(float *X, float *T, long count, long offset, long stride) {
    if (!les || !les->isDSM || !les->C)
        throw Error(MC3D_ECONFIG, "sampling needs LESMODEL DSM with DSMCOEFFRANGE TRUE");
    double *keep[4] = {pr.p, pr.T, pr.mu, pr.sos};
    for (int k = 0; k < 4; ++k)
        hip_check(hipMemcpyAsync(tmp[k], keep[k], sizeof(double) * N, hipMemcpyDeviceToDevice, st), "save state");
    const int saved = rkStep;
    rkStep = 1;  // the current solution is the *1 register set
    try {
        gradientsAndSGS();
    } catch (...) {
        rkStep = saved;
        throw;
    }
    rkStep = saved;
    SampleArgs sa{};
    for (int d = 0; d < 3; ++d) { sa.dU[d] = dU[d]; sa.dV[d] = dV[d]; sa.dW[d] = dW[d]; }
    sa.m = msh;
    sa.C = les->C;
    sa.Prt = les->Prt;
    sa.vol = dom.dx * dom.dy * dom.dz;
    if (opt.musgsAvgType != Options::NONE && !opt.closurePointFeatures) {
        // the features averaged as the DSM averaged its contractions, in DSM work fields
        DSMSGS *d = static_cast<DSMSGS *>(les.get());
        FeatureArgs fa{};
        for (int k = 0; k < 3; ++k) { fa.dU[k] = dU[k]; fa.dV[k] = dV[k]; fa.dW[k] = dW[k]; }
        fa.m = msh;
        fa.vol = sa.vol;
        for (int k = 0; k < kMlpIn; ++k) fa.feat[k] = d->work[1 + k];
        hip_check(launch_closure_features(fa, N, st), "closure features");
        for (int k = 0; k < kMlpIn; ++k) les->doAveraging(fa.feat[k], fa.feat[k]);
        for (int k = 0; k < kMlpIn; ++k) sa.feat[k] = fa.feat[k];
    }
    sa.offset = offset; sa.stride = stride; sa.count = count;
    sa.X = X; sa.T = T;
    hip_check(launch_closure_samples(sa, st), "closure samples");
    for (int k = 0; k < 4; ++k)
        hip_check(hipMemcpyAsync(keep[k], tmp[k], sizeof(double) * N, hipMemcpyDeviceToDevice, st), "restore state");
    hip_check(hipStreamSynchronize(st), "closure samples");
}

void CurvilinearCSolver::hitStats(double out[11]) {
    AbstractDerivatives *dv[3] = {derivX.get(), derivY.get(), derivZ.get()};
    for (int d = 0; d < 3; ++d) {
        const double *in[3] = {pr.U, pr.V, pr.W};
        double *o[3] = {dU[d], dV[d], dW[d]};
        dv[d]->calc1stDerivFields(3, in, o, false);
    }
    const int nb = stats_blocks(N);
    std::vector<double> part((size_t)nb * kStatSums), part2((size_t)nb);
    void *mem = nullptr;
    hip_check(hipMalloc(&mem, sizeof(double) * (size_t)nb * (kStatSums + 1)), "statistics scratch");
    struct Free {
        void *p;
        hipStream_t s;
        ~Free() {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    } guard{mem, st};
    StatsArgs a{};
    for (int d = 0; d < 3; ++d) { a.dU[d] = dU[d]; a.dV[d] = dV[d]; a.dW[d] = dW[d]; }
    a.m = msh;
    a.rho = r1.q[0]; a.U = pr.U; a.V = pr.V; a.W = pr.W; a.mu = pr.mu; a.sos = pr.sos;
    a.partial = static_cast<double *>(mem);
    a.partial2 = a.partial + (size_t)nb * kStatSums;
    hip_check(launch_hit_stats(a, N, st), "statistics");
    hip_check(hipMemcpyAsync(part.data(), a.partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st),
              "statistics");
    StepScalars h;
    hip_check(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st), "statistics");
    hip_check(hipStreamSynchronize(st), "statistics");
    double m[kStatSums] = {};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < kStatSums; ++k) m[k] += part[(size_t)b * kStatSums + k];
    const double Nall = (double)N;
    for (double &v : m) v /= Nall;
    hip_check(launch_hit_stats_rho(r1.q[0], m[0], a.partial2, N, st), "statistics");
    hip_check(hipMemcpyAsync(part2.data(), a.partial2, sizeof(double) * nb, hipMemcpyDeviceToHost, st), "statistics");
    hip_check(hipStreamSynchronize(st), "statistics");
    double rhovar = 0.0;
    for (int b = 0; b < nb; ++b) rhovar += part2[(size_t)b];
    rhovar /= Nall;
    // :521-543
    const double meanRho = m[0], meanMu = m[1], meanSOS = m[6], meanKE = m[8];
    const double meanNu = meanMu / meanRho;
    const double meanTurbDiss = m[2] / meanRho;
    const double uprime = std::sqrt(m[3]), uiprime = std::sqrt(m[4]), urms = std::sqrt(m[5]);
    const double turbMach = urms / meanSOS;
    const double taylorMicro2 = uprime / uiprime;
    const double taylorReyn3 = uprime * taylorMicro2 / meanNu;
    const double k0 = 8.0;
    const double tau = std::sqrt(2 * M_PI) / k0 / uprime;
    const double kolNu = std::pow(std::pow(meanNu, 3.0) / meanTurbDiss, 0.25);
    const double v[11] = {h.time, tau, taylorReyn3, turbMach, meanTurbDiss, uprime, kolNu,
                          std::sqrt(rhovar), std::sqrt(m[9]), std::sqrt(m[10]), meanKE};
    for (int k = 0; k < 11; ++k) out[k] = v[k];
}

void CurvilinearCSolver::preSubStep() {
    preStepBCHandling();
    gradientsAndSGS();
    flux();
}

// src/CurvilinearCSolver_Core.cpp:500-594: velocity and temperature
// derivatives and the SGS fields of the current stage's state
void CurvilinearCSolver::gradientsAndSGS() {
    const Cons &s = state();
    AbstractDerivatives *dv[3] = {derivX.get(), derivY.get(), derivZ.get()};
    if (les && les->isDSM) {
        // DSM changes p, T, mu and sos before the temperature gradient (:568-594)
        for (int d = 0; d < 3; ++d) {
            const double *in[3] = {pr.U, pr.V, pr.W};
            double *out[3] = {dU[d], dV[d], dW[d]};
            dv[d]->calc1stDerivFields(3, in, out, false);
        }
        les->calcMuSGSTaukk(dU, dV, dW, s.q[0], s.q[1], s.q[2], s.q[3]);
        timed("k_dsm_taukk_eos", 7.0,
              [&] { hip_check(launch_dsm_taukk_eos(s.q[0], les->taukk, pr, ig, N, st), "dsm tau_kk"); });
        for (int d = 0; d < 3; ++d) {
            const double *in[1] = {pr.T};
            double *out[1] = {dT[d]};
            dv[d]->calc1stDerivFields(1, in, out, false);
        }
        les->calcKSGS(dT, s.q[0], s.q[1], s.q[2], s.q[3], pr.T);
    } else {
        for (int d = 0; d < 3; ++d) {
            const double *in[4] = {pr.U, pr.V, pr.W, pr.T};
            double *out[4] = {dU[d], dV[d], dW[d], dT[d]};
            dv[d]->calc1stDerivFields(4, in, out, false);
        }
        if (les) les->getSGSViscosity(dU, dV, dW, s.q[0]);
    }
}

// src/CurvilinearCSolver_Core.cpp:615-773: fluxes and their divergence
void CurvilinearCSolver::flux() {
    const Cons &s = state();
    FluxArgs fa{};
    for (int d = 0; d < 3; ++d) { fa.dU[d] = dU[d]; fa.dV[d] = dV[d]; fa.dW[d] = dW[d]; fa.dT[d] = dT[d]; }
    fa.m = msh;
    fa.U = pr.U; fa.V = pr.V; fa.W = pr.W; fa.p = pr.p; fa.mu = pr.mu;
    fa.mu_sgs = les ? les->mu_sgs : nullptr;
    fa.k_sgs = les ? les->k_sgs : nullptr;
    fa.rho = s.q[0]; fa.rhoU = s.q[1]; fa.rhoV = s.q[2]; fa.rhoW = s.q[3]; fa.rhoE = s.q[4];
    for (int e = 0; e < 15; ++e) fa.pre[e] = pre[e];
    // 1: mu_sgs with k_sgs from Pr_t (Vreman); 2: mu_sgs and k_sgs fields (DSM, learned)
    fa.les = opt.lesModel == Options::NOMODEL ? 0 : opt.lesModel == Options::VREMAN ? 1 : 2;
    fa.Pr_t = Pr_t;
    // read: 12 derivatives, 10 metrics, U V W p mu, 5 conserved (+ mu_sgs, k_sgs); written: 15
    const double flux_fields = 32.0 + 15.0 + (fa.les ? 1.0 : 0.0) + (fa.les == 2 ? 1.0 : 0.0);
    timed("k_flux", flux_fields, [&] { hip_check(launch_flux(fa, ig, N, st), "flux"); });
    // divergence: xi of the _1 fluxes, then eta of _2 and zeta of _3 added in
    // (src/CurvilinearCSolver_Core.cpp:763-773 with the sums of :810,839,...)
    const double *p1[5] = {pre[0], pre[1], pre[2], pre[3], pre[4]};
    const double *p2[5] = {pre[5], pre[6], pre[7], pre[8], pre[9]};
    const double *p3[5] = {pre[10], pre[11], pre[12], pre[13], pre[14]};
    derivX->calc1stDerivFields(5, p1, div, false);
    derivY->calc1stDerivFields(5, p2, div, true);
    derivZ->calc1stDerivFields(5, p3, div, true);
    // the wall rows keep the continuity flux of their own direction alone (:702-760)
    AbstractDerivatives *dv[3] = {derivX.get(), derivY.get(), derivZ.get()};
    for (int d = 0; d < 3; ++d)
        if (cont_[d]) {
            const double *in[1] = {pre[5 * d]};
            double *out[1] = {cont_[d]};
            dv[d]->calc1stDerivFields(1, in, out, false);
        }
}

void CurvilinearCSolver::stage_update(int op, double a, double b, double c) {
    RkArgs ra{};
    ra.r1 = r1; ra.rk = rk_; ra.k2 = k2; ra.r2 = r2;
    for (int e = 0; e < 5; ++e) ra.div[e] = div[e];
    ra.J = msh.J;
    ra.sc = sc;
    ra.member_pts = dom.members > 1 ? (unsigned)dom.gN : 0u;
    ra.op = op;
    ra.a = a; ra.b = b; ra.c = c;
    // the non-conserved update of a stage that writes rk and is not followed by
    // a filter rides along (its result does not depend on where it is computed)
    const bool writes_rk = op != kTvd3 && op != kRk4Last && op != kLsLast;
    ra.eos = (writes_rk && !rkLast && !opt.subStepFiltering) ? 1 : 0;
    ra.pr = pr;
    ra.g = ig;
    ra.bc = rkBoundaryArgs();
    eos_done_ = ra.eos != 0;
    // fields: div (5) + J + previous momentum k2 (3) read, k2 (5) written, plus the
    // registers, plus U V W p T mu sos written when the update rides along
    static const double reg[9] = {10, 15, 15, 15, 20, 10, 15, 20, 20};
    timed("k_rhs_rk", 14.0 + reg[op] + (ra.eos ? 7.0 : 0.0),
          [&] { hip_check(launch_rhs_rk(ra, N, st), "stage update"); });
}

// src/CurvilinearCSolver.hpp:346-353
void CurvilinearCSolver::updateData() {
    if (rkLast || opt.subStepFiltering) filterConservedData();
    updateNonConservedData();
}

// src/CurvilinearCSolver_Core.cpp:949-1155
void CurvilinearCSolver::filterConservedData() {
    const Cons &in = rkLast ? r2 : rk_;
    const Cons &out = rkLast ? r1 : rk_;
    if (timeStep % opt.filterStep == 0) {
        filterTimeStep++;
        AbstractFilter *f[3] = {filtX.get(), filtY.get(), filtZ.get()};
        int order[3];
        if (filterTimeStep % 3 == 1) { order[0] = 1; order[1] = 2; order[2] = 0; }       // Y, Z, X
        else if (filterTimeStep % 3 == 2) { order[0] = 2; order[1] = 0; order[2] = 1; }  // Z, X, Y
        else { order[0] = 0; order[1] = 1; order[2] = 2; }                                // X, Y, Z
        const double *src[5];
        double *t[5], *dst[5];
        for (int q = 0; q < 5; ++q) { src[q] = in.q[q]; t[q] = tmp[q]; dst[q] = out.q[q]; }
        const double *tc[5] = {t[0], t[1], t[2], t[3], t[4]};
        f[order[0]]->filterFields(5, src, t);
        f[order[1]]->filterFields(5, tc, t);
        f[order[2]]->filterFields(5, tc, dst);
    } else if (rkLast) {
        for (int q = 0; q < 5; ++q)
            hip_check(hipMemcpyAsync(r1.q[q], r2.q[q], sizeof(double) * N, hipMemcpyDeviceToDevice, st), "copy");
    }
}

void CurvilinearCSolver::updateNonConservedData() {
    if (eos_done_) {  // already done by the stage kernel
        eos_done_ = false;
        return;
    }
    // read: 5 conserved; written: U V W p T mu sos
    timed("k_eos", 12.0, [&] { hip_check(launch_eos(rkLast ? r1 : rk_, pr, ig, N, st), "eos"); });
}

// src/Utils.cpp:415-431: the sequential fold std::min(phi, acc) from 1e6 (and
// std::max from -1e6); for finite data the block-wise min/max gives the same
void CurvilinearCSolver::fieldRange(const double *phi, double &lo, double &hi) {
    const int nb = 512;  // 2 * nb doubles fit the scratch buffer
    hip_check(launch_minmax_partial(phi, cfl_partial, nb, N, st), "range");
    std::vector<double> part(2 * nb);
    hip_check(hipMemcpyAsync(part.data(), cfl_partial, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st),
              "range read-back");
    hip_check(hipStreamSynchronize(st), "range read-back");
    lo = 1000000.0;
    hi = -1000000.0;
    for (int b = 0; b < nb; ++b) {
        lo = std::min(part[2 * b], lo);
        hi = std::max(part[2 * b + 1], hi);
    }
}

void CurvilinearCSolver::getRange(const double *phi, const char *name) {
    double lo, hi;
    fieldRange(phi, lo, hi);
    std::printf("   Range of %s: %g:%g\n", name, lo, hi);
}

// src/AbstractSingleBlockMesh.cpp:1336-1393: I_k = (dJ_1k/dxi + dJ_2k/deta) +
// dJ_3k/dzeta with the solver's plain first derivatives (no periodic
// transformation, as in the reference), then the ranges
void CurvilinearCSolver::metricIdentity(double range[6], bool print) {
    void *mem = nullptr;
    hip_check(hipMalloc(&mem, sizeof(double) * (size_t)N), "identity scratch");
    struct Free {
        void *p;
        hipStream_t s;
        ~Free() {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    } guard{mem, st};
    double *sum = static_cast<double *>(mem);
    if (print) std::printf(" > Checking the values of the metric indentities, values should be small\n");
    for (int k = 0; k < 3; ++k) {
        const double *a = msh_rw(1 + k), *b = msh_rw(4 + k), *c = msh_rw(7 + k);
        derivX->calc1stDerivFields(1, &a, &sum, false);
        derivY->calc1stDerivFields(1, &b, &sum, true);
        derivZ->calc1stDerivFields(1, &c, &sum, true);
        fieldRange(sum, range[2 * k], range[2 * k + 1]);
        if (print) std::printf("   Range of Metric Identity %d: %g:%g\n", k + 1, range[2 * k], range[2 * k + 1]);
    }
    if (print) std::fflush(stdout);
}

// src/CurvilinearCSolver_Core.cpp:1244-1329
void CurvilinearCSolver::checkSolution() {
    if (timeStep % opt.checkStep != 0) return;
    StepScalars h;
    hip_check(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st), "scalar read");
    hip_check(hipStreamSynchronize(st), "scalar read");
    const auto now = std::chrono::steady_clock::now();
    const double since = std::chrono::duration<double>(now - last_check_).count();
    last_check_ = now;
    std::printf("\n-------------------------------------------------\n");
    std::printf(" Step = %d, time = %g, dt = %g\n", timeStep, h.time, h.dt);
    std::printf("-------------------------------------------------\n");
    std::printf("  Time since last timestep = %g\n", since);
    hip_check(launch_contravariant(pr, msh, N, st), "contravariant velocities");
    getRange(r1.q[0], "RHO");
    getRange(pr.U, "U");
    getRange(pr.V, "V");
    getRange(pr.W, "W");
    getRange(pr.Ucurv, "Ucurv");
    getRange(pr.Vcurv, "Vcurv");
    getRange(pr.Wcurv, "Wcurv");
    getRange(pr.p, "P");
    getRange(pr.T, "T");
    getRange(pr.mu, "mu");
    getRange(r1.q[4], "RHOE");
    getRange(pr.sos, "SOS");
    if (les) {
        getRange(les->mu_sgs, "MU SGS");
        if (les->isDSM) {
            getRange(les->taukk, "TAUKK ");
            getRange(les->k_sgs, "K SGS");
            if (les->C) {
                getRange(les->C, "LES C");
                getRange(les->CI, "LES CI");
                getRange(les->Prt, "LES Prt");
            }
        }
    }
    // src/CurvilinearCSolver_Core.cpp:1295-1316
    if (stats) {
        if (stats->velocityStats)
            for (const char *f : {"UAVG", "URMS", "VAVG", "VRMS", "WAVG", "WRMS", "UVREY", "UWREY", "VWREY"})
                getRange(stats->field(f), f);
        if (stats->thermoStats)
            for (const char *f : {"RHOAVG", "RHORMS", "PAVG", "PRMS", "TAVG", "TRMS"}) getRange(stats->field(f), f);
    }
    std::fflush(stdout);
}

void CurvilinearCSolver::set_turbdata(const std::string &path, int every) {
    if (every < 1) throw Error(MC3D_EINVAL, "the statistics interval must be positive");
    if (members() > 1) throw Error(MC3D_ECONFIG, "stacked members keep no HIT statistics");
    turb_path_ = path;
    turb_every_ = every;
}

void CurvilinearCSolver::postStep() {
    // fullStepTemporalHook of the HIT driver (tests/hit/Samtaney_D4/MPI_3DCompact.cpp:372-557,
    // called first in postStep, src/CurvilinearCSolver.hpp:355-357): every 25
    // steps one line of 11 statistics, precision 17, appended to turbdata.out
    if (!turb_path_.empty() && timeStep % turb_every_ == 0) {
        double v[11];
        hitStats(v);
        FILE *f = std::fopen(turb_path_.c_str(), "a");
        if (!f) throw Error(MC3D_EINVAL, "cannot append to " + turb_path_);
        bool ok = true;
        for (int k = 0; k < 11; ++k) ok = std::fprintf(f, "%.17g ", v[k]) > 0 && ok;
        ok = std::fputc('\n', f) != EOF && ok;
        ok = std::fclose(f) == 0 && ok;
        if (!ok) throw Error(MC3D_EINVAL, "short write to " + turb_path_);
    }
    checkSolution();
    // src/CurvilinearCSolver.hpp:362-364
    if (stats && timeStep % opt.stats_interval == 0) {
        StepScalars h;
        hip_check(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st), "scalar read");
        hip_check(hipStreamSynchronize(st), "scalar read");
        stats->updateStatsFields(h.time);
    }
    const double *mesh[3] = {mesh_[0].empty() ? nullptr : mesh_[0].data(),
                             mesh_[1].empty() ? nullptr : mesh_[1].data(),
                             mesh_[2].empty() ? nullptr : mesh_[2].data()};
    // src/CurvilinearCSolver.hpp:367-369
    if (dumps_ && timeStep % opt.dumpStep == 0) dumpSolution(std::string(), mesh);
    writeImages();
    // src/CurvilinearCSolver_Core.cpp:1761-1796: a file named killsolver in the
    // working directory ends the run (and is deleted)
    if (FILE *kf = std::fopen("killsolver", "r")) {
        std::fclose(kf);
        if (std::remove("killsolver") != 0) std::printf(" > ERROR DELETING KILLFILE! \n");
        std::printf("===================\n FOUND KILL SOLVER \n===================\n");
        std::fflush(stdout);
        endFlag = true;
    }
    if (timeStep >= opt.maxTimeStep) endFlag = true;
    if (!endFlag) {
        StepScalars h;
        hip_check(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st), "scalar read");
        hip_check(hipStreamSynchronize(st), "scalar read");
        if (h.time >= opt.maxTime) endFlag = true;
    }
    // checkEnd dumps the final state, src/CurvilinearCSolver_Core.cpp:1822
    if (dumps_ && endFlag) dumpSolution(std::string(), mesh);
}

double *CurvilinearCSolver::field_by_name(const std::string &n) {
    static const char *cons[5] = {"rho", "rhoU", "rhoV", "rhoW", "rhoE"};
    for (int q = 0; q < 5; ++q) {
        if (n == cons[q]) return r1.q[q];
        if (n == std::string(cons[q]) + "k") return rk_.q[q];
        if (n == std::string(cons[q]) + "k2") return k2.q[q];
        if (n == std::string(cons[q]) + "2") return r2.q[q];
    }
    if (n == "U") return pr.U;
    if (n == "V") return pr.V;
    if (n == "W") return pr.W;
    if (n == "p") return pr.p;
    if (n == "T") return pr.T;
    if (n == "mu") return pr.mu;
    if (n == "sos") return pr.sos;
    if (n == "Ucurv" || n == "Vcurv" || n == "Wcurv") {
        hip_check(launch_contravariant(pr, msh, N, st), "contravariant velocities");
        return n == "Ucurv" ? pr.Ucurv : n == "Vcurv" ? pr.Vcurv : pr.Wcurv;
    }
    if (n == "mu_sgs") return les ? les->mu_sgs : nullptr;
    if (n == "k_sgs") return les ? les->k_sgs : nullptr;
    if (n == "taukk") return les ? les->taukk : nullptr;
    if (n == "C") return les ? les->C : nullptr;
    if (n == "CI") return les ? les->CI : nullptr;
    if (n == "Prt") return les ? les->Prt : nullptr;
    // computational derivatives of the last evaluated stage: dU0 = dU/dxi, dU1 = dU/deta, ...
    if (n.size() == 3 && n[0] == 'd' && (n[1] == 'U' || n[1] == 'V' || n[1] == 'W' || n[1] == 'T') &&
        n[2] >= '0' && n[2] <= '2') {
        const int d = n[2] - '0';
        return n[1] == 'U' ? dU[d] : n[1] == 'V' ? dV[d] : n[1] == 'W' ? dW[d] : dT[d];
    }
    if (n == "sigma") return sigma_;
    if (stats)
        if (double *p = stats->field(n)) return p;
    if (n.size() > 3 && n.compare(n.size() - 3, 3, "Avg") == 0)  // sponge averages rhoAvg .. rhoEAvg
        for (int q = 0; q < 5; ++q)
            if (n == std::string(cons[q]) + "Avg") return spgAvg_[q];
    static const char *met[10] = {"J", "J11", "J12", "J13", "J21", "J22", "J23", "J31", "J32", "J33"};
    for (int k = 0; k < 10; ++k)
        if (n == met[k]) return Jm[k];
    return nullptr;
}

}  // namespace mc3d
