// C ABI of the SGS models over caller-owned device fields (include/mc3d.h):
// the reference's AbstractSGS interface (src/AbstractSGS.hpp:35-41) without
// the solver around it.  Exceptions never cross it.
#include <hip/hip_runtime.h>

#include <string>

#include "../capi_common.h"
#include "capi_guard.hpp"
#include "closure.hpp"
#include "mc3d.h"
#include "solver.hpp"

struct mc3d_sgs {
    mc3d::CurvilinearCSolver *cs = nullptr;  // SGS-only mode: the model and its work fields
    bool dsm_ran = false;                    // calcKSGS reads fields calcMuSGSTaukk filtered
};

using mc3d::guarded;

namespace {

int model_check(const mc3d_sgs *s, mc3d::Options::LESModel m, const char *what) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null SGS handle");
    if (s->cs->opt.lesModel != m) return mc3d::fail(MC3D_EINVAL, std::string(what) + " of another LESMODEL");
    return MC3D_OK;
}

bool all(const double *const *p, int n) {
    if (!p) return false;
    for (int k = 0; k < n; ++k)
        if (!p[k]) return false;
    return true;
}

// the Cartesian gradient rows as the model classes take them: d(u_i)[j] = gradU[3 i + j]
void rows(const double *const gradU[9], const double *r[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = gradU[3 * i + j];
}

}  // namespace

extern "C" {

int mc3d_sgs_create(const char *options_text, int layout, const double *J, void *stream, mc3d_sgs **out) {
    if (!out) return mc3d::fail(MC3D_EINVAL, "null output handle");
    *out = nullptr;
    if (!options_text || !J) return mc3d::fail(MC3D_EINVAL, "null argument");
    mc3d::Options o;
    std::string err;
    int rc = mc3d::Options::parse(options_text, o, err);
    if (rc) return mc3d::fail(rc, "options: " + err);
    auto *s = new mc3d_sgs();
    rc = guarded([&] {
        s->cs = new mc3d::CurvilinearCSolver(o, static_cast<hipStream_t>(stream),
                                             mc3d::CurvilinearCSolver::SgsOnly{layout, J});
    });
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return MC3D_OK;
}

void mc3d_sgs_destroy(mc3d_sgs *s) {
    if (!s) return;
    try {
        delete s->cs;
    } catch (...) {
    }
    delete s;
}

int mc3d_sgs_vreman(mc3d_sgs *s, const double *const gradU[9], const double *rho, double *mu_sgs) {
    int rc = model_check(s, mc3d::Options::VREMAN, "mc3d_sgs_vreman");
    if (rc) return rc;
    if (!all(gradU, 9) || !rho || !mu_sgs) return mc3d::fail(MC3D_EINVAL, "null field");
    return guarded([&] {
        const double *r[3][3];
        rows(gradU, r);
        mc3d::AbstractSGS *les = s->cs->les.get();
        les->mu_sgs = mu_sgs;
        les->getSGSViscosity(r[0], r[1], r[2], rho);
    });
}

int mc3d_sgs_dsm_mu_taukk(mc3d_sgs *s, const double *const gradU[9], const double *rho, const double *rhoU,
                          const double *rhoV, const double *rhoW, double *mu_sgs, double *taukk) {
    int rc = model_check(s, mc3d::Options::DSM, "mc3d_sgs_dsm_mu_taukk");
    if (rc) return rc;
    if (!all(gradU, 9) || !rho || !rhoU || !rhoV || !rhoW || !mu_sgs || !taukk)
        return mc3d::fail(MC3D_EINVAL, "null field");
    return guarded([&] {
        const double *r[3][3];
        rows(gradU, r);
        mc3d::AbstractSGS *les = s->cs->les.get();
        les->mu_sgs = mu_sgs;
        les->taukk = taukk;
        les->calcMuSGSTaukk(r[0], r[1], r[2], rho, rhoU, rhoV, rhoW);
        s->dsm_ran = true;
    });
}

int mc3d_sgs_dsm_ksgs(mc3d_sgs *s, const double *const gradT[3], const double *rho, const double *rhoU,
                      const double *rhoV, const double *rhoW, const double *T, double *k_sgs) {
    int rc = model_check(s, mc3d::Options::DSM, "mc3d_sgs_dsm_ksgs");
    if (rc) return rc;
    if (!all(gradT, 3) || !rho || !rhoU || !rhoV || !rhoW || !T || !k_sgs)
        return mc3d::fail(MC3D_EINVAL, "null field");
    if (!s->dsm_ran) return mc3d::fail(MC3D_EINVAL, "calcKSGS needs mc3d_sgs_dsm_mu_taukk of the same state first");
    return guarded([&] {
        mc3d::AbstractSGS *les = s->cs->les.get();
        les->k_sgs = k_sgs;
        les->calcKSGS(gradT, rho, rhoU, rhoV, rhoW, T);
    });
}

int mc3d_sgs_coefficients(mc3d_sgs *s, double *C, double *CI, double *Prt) {
    int rc = model_check(s, mc3d::Options::DSM, "mc3d_sgs_coefficients");
    if (rc) return rc;
    const mc3d::AbstractSGS *les = s->cs->les.get();
    if (!les->C) return mc3d::fail(MC3D_ECONFIG, "the coefficients need DSMCOEFFRANGE TRUE");
    return guarded([&] {
        const size_t bytes = sizeof(double) * (size_t)s->cs->N;
        const double *src[3] = {les->C, les->CI, les->Prt};
        double *dst[3] = {C, CI, Prt};
        for (int k = 0; k < 3; ++k)
            if (dst[k])
                mc3d::hip_check(hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToDevice, s->cs->st),
                                "coefficient copy");
    });
}

int mc3d_sgs_set_closure(mc3d_sgs *s, const mc3d_closure *c) {
    int rc = model_check(s, mc3d::Options::LEARNED, "mc3d_sgs_set_closure");
    if (rc) return rc;
    if (!c) return mc3d::fail(MC3D_EINVAL, "null closure");
    return guarded([&] { static_cast<mc3d::LearnedSGS *>(s->cs->les.get())->set_network(*mc3d_closure_net(c)); });
}

int mc3d_sgs_learned(mc3d_sgs *s, const double *const gradU[9], const double *const gradT[3], const double *rho,
                     const double *rhoU, const double *rhoV, const double *rhoW, const double *T, double *mu_sgs,
                     double *k_sgs) {
    int rc = model_check(s, mc3d::Options::LEARNED, "mc3d_sgs_learned");
    if (rc) return rc;
    if (!all(gradU, 9) || !rho || !mu_sgs || !k_sgs) return mc3d::fail(MC3D_EINVAL, "null field");
    mc3d::CurvilinearCSolver *cs = s->cs;
    const bool fallback = cs->opt.closureFallback;
    if (fallback && (!all(gradT, 3) || !rhoU || !rhoV || !rhoW || !T))
        return mc3d::fail(MC3D_EINVAL, "the DSM fallback needs gradT, rhoU, rhoV, rhoW and T");
    return guarded([&] {
        const double *r[3][3];
        rows(gradU, r);
        // the fallback DSM reads the stage state and the temperature gradient
        cs->rkStep = 1;
        cs->r1.q[0] = const_cast<double *>(rho);
        cs->r1.q[1] = const_cast<double *>(rhoU);
        cs->r1.q[2] = const_cast<double *>(rhoV);
        cs->r1.q[3] = const_cast<double *>(rhoW);
        for (int d = 0; d < 3; ++d) cs->dT[d] = fallback ? const_cast<double *>(gradT[d]) : nullptr;
        cs->pr.T = const_cast<double *>(T);
        mc3d::AbstractSGS *les = cs->les.get();
        les->mu_sgs = mu_sgs;
        les->k_sgs = k_sgs;
        les->getSGSViscosity(r[0], r[1], r[2], rho);
    });
}

int mc3d_sgs_closure_stats(mc3d_sgs *s, long *evaluations, long *outside_points, long *fallback_stages,
                           long *fallback_points) {
    int rc = model_check(s, mc3d::Options::LEARNED, "mc3d_sgs_closure_stats");
    if (rc) return rc;
    const auto *l = static_cast<const mc3d::LearnedSGS *>(s->cs->les.get());
    if (evaluations) *evaluations = l->evaluations;
    if (outside_points) *outside_points = l->outside_points;
    if (fallback_stages) *fallback_stages = l->fallback_stages;
    if (fallback_points) *fallback_points = l->fallback_points;
    return MC3D_OK;
}

}  // extern "C"
