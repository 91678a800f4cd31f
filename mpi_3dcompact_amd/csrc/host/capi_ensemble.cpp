// C ABI of the ensembles (include/mc3d.h).  Exceptions never cross it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "../capi_common.h"
#include "capi_guard.hpp"
#include "ensemble.hpp"
#include "mc3d.h"

struct mc3d_ensemble {
    mc3d::Ensemble *e = nullptr;
};

using mc3d::guarded;

extern "C" {

int mc3d_ensemble_create(const char *options_text, int members, const int *dims, const double *mu_ref,
                         mc3d_ensemble **out) {
    if (!out) return mc3d::fail(MC3D_EINVAL, "null output handle");
    *out = nullptr;
    if (!options_text || !dims) return mc3d::fail(MC3D_EINVAL, "null argument");
    mc3d::Options o;
    std::string err;
    int rc = mc3d::Options::parse(options_text, o, err);
    if (rc) return mc3d::fail(rc, "options: " + err);
    auto *h = new mc3d_ensemble();
    rc = guarded([&] { h->e = new mc3d::Ensemble(o, members, dims, mu_ref); });
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return MC3D_OK;
}

int mc3d_ensemble_set_metrics(mc3d_ensemble *h, int member, const double *const metrics[10]) {
    if (!h || !metrics) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { h->e->setMetrics(member, metrics); });
}

int mc3d_ensemble_set_initial(mc3d_ensemble *h, int member, const double *rho0, const double *U0,
                              const double *V0, const double *W0, const double *p0) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    return guarded([&] { h->e->setInitial(member, rho0, U0, V0, W0, p0); });
}

int mc3d_ensemble_tune(mc3d_ensemble *h, const char *cache_path) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    return guarded([&] { h->e->tune(cache_path ? std::string(cache_path) : std::string()); });
}

int mc3d_ensemble_step(mc3d_ensemble *h, int nsteps) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    return guarded([&] { h->e->step(nsteps); });
}

int mc3d_ensemble_use_graphs(mc3d_ensemble *h, int on) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    h->e->useGraphs(on != 0);
    return MC3D_OK;
}

int mc3d_ensemble_get_field(mc3d_ensemble *h, int member, const char *name, double *host_out) {
    if (!h || !name) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { h->e->getField(member, name, host_out); });
}

int mc3d_ensemble_scalars(mc3d_ensemble *h, int member, double *time, double *dt, int *timestep) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    return guarded([&] {
        const mc3d::StepScalars s = h->e->scalars(member);
        if (time) *time = s.time;
        if (dt) *dt = s.dt;
        if (timestep) *timestep = h->e->timeStep();
    });
}

int mc3d_ensemble_stream(mc3d_ensemble *h, void **stream) {
    if (!h || !stream) return mc3d::fail(MC3D_EINVAL, "null argument");
    *stream = h->e->stream();
    return MC3D_OK;
}

int mc3d_ensemble_sync(mc3d_ensemble *h) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    return guarded([&] { h->e->sync(); });
}

int mc3d_ensemble_info(mc3d_ensemble *h, char *buf, int buflen, int *needed) {
    if (!h) return mc3d::fail(MC3D_EINVAL, "null ensemble");
    std::string j;
    const int rc = guarded([&] { j = h->e->info(); });
    if (rc) return rc;
    if (needed) *needed = (int)j.size() + 1;
    if (!buf || buflen < (int)j.size() + 1) return mc3d::fail(MC3D_EINVAL, "report buffer too small");
    std::memcpy(buf, j.c_str(), j.size() + 1);
    return MC3D_OK;
}

void mc3d_ensemble_destroy(mc3d_ensemble *h) {
    if (!h) return;
    delete h->e;
    delete h;
}

}  // extern "C"
