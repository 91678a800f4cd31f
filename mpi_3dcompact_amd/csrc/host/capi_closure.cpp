// C ABI of the learned SGS closure network (include/mc3d.h).
#include <hip/hip_runtime.h>

#include <cmath>

#include "capi_guard.hpp"
#include "closure.hpp"

struct mc3d_closure {
    mc3d::Closure *net = nullptr;
};

using mc3d::guarded;

const mc3d::Closure *mc3d_closure_net(const mc3d_closure *c) { return c->net; }

namespace {

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int mc3d_closure_create(unsigned long long seed, mc3d_closure **out) {
    if (!out) return mc3d::fail(MC3D_EINVAL, "null argument");
    *out = nullptr;
    auto *c = new (std::nothrow) mc3d_closure();
    if (!c) return mc3d::fail(MC3D_ENOMEM, "host allocation failed");
    const int rc = guarded([&] { c->net = new mc3d::Closure(seed); });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return MC3D_OK;
}

void mc3d_closure_destroy(mc3d_closure *c) {
    if (!c) return;
    try {
        delete c->net;
    } catch (...) {
    }
    delete c;
}

int mc3d_closure_set_params(mc3d_closure *c, const float *params) {
    if (!c || !params) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { c->net->set_params_host(params); });
}

int mc3d_closure_get_params(const mc3d_closure *c, float *params) {
    if (!c || !params) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { c->net->get_params_host(params); });
}

int mc3d_closure_set_scaling(mc3d_closure *c, const float in_mean[7], const float in_std[7],
                             const float out_mean[2], const float out_std[2], const float lo[7],
                             const float hi[7]) {
    if (!c || !in_mean || !in_std || !out_mean || !out_std || !lo || !hi)
        return mc3d::fail(MC3D_EINVAL, "null argument");
    mc3d::MlpScaling s{};
    for (int k = 0; k < mc3d::kMlpIn; ++k) {
        if (!(in_std[k] > 0.0f) || !std::isfinite(in_mean[k]) || !(lo[k] <= hi[k]))
            return mc3d::fail(MC3D_EINVAL, "input scaling needs finite means, positive std and lo <= hi");
        s.in_mean[k] = in_mean[k]; s.in_std[k] = in_std[k]; s.lo[k] = lo[k]; s.hi[k] = hi[k];
    }
    for (int k = 0; k < mc3d::kMlpOut; ++k) {
        if (!(out_std[k] > 0.0f) || !std::isfinite(out_mean[k]))
            return mc3d::fail(MC3D_EINVAL, "output scaling needs finite means and positive std");
        s.out_mean[k] = out_mean[k]; s.out_std[k] = out_std[k];
    }
    c->net->sc = s;
    c->net->scaled = true;
    return MC3D_OK;
}

int mc3d_closure_get_scaling(const mc3d_closure *c, float in_mean[7], float in_std[7], float out_mean[2],
                             float out_std[2], float lo[7], float hi[7]) {
    if (!c || !in_mean || !in_std || !out_mean || !out_std || !lo || !hi)
        return mc3d::fail(MC3D_EINVAL, "null argument");
    const mc3d::MlpScaling &s = c->net->sc;
    for (int k = 0; k < mc3d::kMlpIn; ++k) {
        in_mean[k] = s.in_mean[k]; in_std[k] = s.in_std[k]; lo[k] = s.lo[k]; hi[k] = s.hi[k];
    }
    for (int k = 0; k < mc3d::kMlpOut; ++k) { out_mean[k] = s.out_mean[k]; out_std[k] = s.out_std[k]; }
    return MC3D_OK;
}

int mc3d_closure_gradient(mc3d_closure *c, const float *X, const float *T, long n, float *grad, double *loss,
                          void *stream) {
    if (!c || !X || !T || !grad) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (n <= 0) return mc3d::fail(MC3D_EINVAL, "no samples");
    return guarded([&] {
        const double l = c->net->gradient(X, T, n, grad, as_stream(stream));
        if (loss) *loss = l;
    });
}

int mc3d_closure_adam_step(mc3d_closure *c, const float *grad, float lr, void *stream) {
    if (!c || !grad) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { c->net->adam(grad, lr, as_stream(stream)); });
}

int mc3d_closure_train(mc3d_closure *c, const float *X, const float *T, long n, int epochs, int batch, float lr,
                       unsigned long long seed, double *epoch_losses, void *stream) {
    if (!c || !X || !T) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (n <= 0 || epochs <= 0 || batch <= 0) return mc3d::fail(MC3D_EINVAL, "empty training problem");
    return guarded([&] { c->net->train(X, T, n, epochs, batch, lr, seed, epoch_losses, as_stream(stream)); });
}

int mc3d_closure_infer(const mc3d_closure *c, const float *X, float *Y, long n, void *stream) {
    if (!c || (n > 0 && (!X || !Y))) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (n < 0) return mc3d::fail(MC3D_EINVAL, "negative sample count");
    return guarded([&] { c->net->infer(X, Y, n, as_stream(stream)); });
}

int mc3d_closure_save(const mc3d_closure *c, const char *path) {
    if (!c || !path) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { c->net->save(path); });
}

int mc3d_closure_load(mc3d_closure *c, const char *path) {
    if (!c || !path) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { c->net->load(path); });
}

}  // extern "C"
