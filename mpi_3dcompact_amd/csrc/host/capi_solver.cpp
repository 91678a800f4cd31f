// C ABI of the whole solver (include/mc3d.h).  Exceptions never cross it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../capi_common.h"
#include "capi_guard.hpp"
#include "closure.hpp"
#include "dump.hpp"
#include "mesh.hpp"
#include "mc3d.h"
#include "solver.hpp"

struct mc3d_solver {
    hipStream_t stream = nullptr;
    mc3d::CurvilinearCSolver *cs = nullptr;
    bool metrics = false, initial = false, sponge = false;
};

using mc3d::guarded;

extern "C" {

int mc3d_options_check(const char *options_text, char *buf, int buflen, int *needed) {
    if (!options_text) return mc3d::fail(MC3D_EINVAL, "null options text");
    if (buflen < 0 || (buflen > 0 && !buf)) return mc3d::fail(MC3D_EINVAL, "bad output buffer");
    mc3d::Options o;
    std::string err;
    const int rc = mc3d::Options::parse(options_text, o, err);
    if (rc) return mc3d::fail(rc, "options: " + err);
    const std::string d = o.describe();
    if (needed) *needed = (int)d.size() + 1;
    if (buflen > 0) {
        const size_t n = std::min((size_t)buflen - 1, d.size());
        std::memcpy(buf, d.data(), n);
        buf[n] = '\0';
    }
    return MC3D_OK;
}

int mc3d_solver_create(const char *options_text, mc3d_solver **out) {
    if (!out) return mc3d::fail(MC3D_EINVAL, "null output handle");
    *out = nullptr;
    if (!options_text) return mc3d::fail(MC3D_EINVAL, "null options text");
    mc3d::Options o;
    std::string err;
    int rc = mc3d::Options::parse(options_text, o, err);
    if (rc) return mc3d::fail(rc, "options: " + err);
    if (o.fromRestart || o.onlyGridFromRestart) {
        // grid (and, for a restart, step and time) from the file header,
        // src/MPI_3DCompact.cpp:80-113
        rc = guarded([&] {
            const mc3d::DumpHeader h = mc3d::read_dump_header(o.filename);
            o.Nx = h.Nx; o.Ny = h.Ny; o.Nz = h.Nz;
            if (o.fromRestart) { o.timeStep = h.timeStep; o.time = h.time; }
            else { o.timeStep = 0; o.time = 0.0; }
        });
        if (rc) return rc;
        if (o.Nx <= 10 || o.Ny <= 10 || o.Nz <= 10)
            return mc3d::fail(MC3D_ECONFIG, "restart grid: every dimension must exceed 10");
    }
    auto *s = new mc3d_solver();
    rc = guarded([&] {
        mc3d::hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream creation");
        s->cs = new mc3d::CurvilinearCSolver(o, s->stream);
    });
    if (rc) {
        delete s->cs;
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return rc;
    }
    *out = s;
    return MC3D_OK;
}

int mc3d_solver_set_metrics(mc3d_solver *s, const double *const metrics[10]) {
    if (!s || !metrics) return mc3d::fail(MC3D_EINVAL, "null argument");
    for (int k = 0; k < 10; ++k)
        if (!metrics[k]) return mc3d::fail(MC3D_EINVAL, "null metric array");
    return guarded([&] { s->cs->set_metrics_host(metrics); s->metrics = true; });
}

int mc3d_solver_set_mesh(mc3d_solver *s, const double *x, const double *y, const double *z) {
    if (!s || !x || !y || !z) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] {
        s->cs->set_mesh_host(x, y, z);
        s->metrics = true;
        if (s->cs->hasSponge()) s->sponge = true;
    });
}

int mc3d_solver_metric_identity(mc3d_solver *s, double range[6]) {
    if (!s || !range) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (!s->metrics) return mc3d::fail(MC3D_ECONFIG, "no metrics: mc3d_solver_set_mesh or mc3d_solver_set_metrics");
    return guarded([&] { s->cs->metricIdentity(range, false); });
}

int mc3d_mesh_metrics(const int fd[3], const int bc[3], const int dims[3], const double trans[9],
                      const double *x, const double *y, const double *z, double *const out[10],
                      void *stream) {
    if (!fd || !bc || !dims || !x || !y || !z || !out) return mc3d::fail(MC3D_EINVAL, "null argument");
    for (int k = 0; k < 10; ++k)
        if (!out[k]) return mc3d::fail(MC3D_EINVAL, "null metric array");
    for (int d = 0; d < 3; ++d) {
        if (fd[d] != MC3D_CD2 && fd[d] != MC3D_PADE6 && fd[d] != MC3D_PENTA10)
            return mc3d::fail(MC3D_EINVAL, "unknown derivative scheme");
        if (bc[d] != MC3D_PERIODIC && bc[d] != MC3D_DIRICHLET) return mc3d::fail(MC3D_EINVAL, "unknown boundary");
        if (dims[d] < 6) return mc3d::fail(MC3D_EINVAL, "each direction needs at least 6 points");
        if (bc[d] == MC3D_PERIODIC && fd[d] == MC3D_CD2)
            return mc3d::fail(MC3D_ECONFIG, "metrics: CD2 on a periodic direction has no transformed-periodic "
                                            "form (src/AbstractSingleBlockMesh.cpp:94-97)");
    }
    return guarded([&] {
        mc3d::Options o;
        o.Nx = dims[0]; o.Ny = dims[1]; o.Nz = dims[2];
        mc3d::Options::BCType *bt[3] = {&o.bcXType, &o.bcYType, &o.bcZType};
        for (int d = 0; d < 3; ++d)
            *bt[d] = bc[d] == MC3D_PERIODIC ? mc3d::Options::PERIODIC_SOLVE : mc3d::Options::DIRICHLET_SOLVE;
        const mc3d::Domain dom(o);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        using AD = mc3d::AbstractDerivatives;
        mc3d::LineDerivative dx(fd[0], dom, o.bcXType, AD::DIRX, st), dy(fd[1], dom, o.bcYType, AD::DIRY, st),
            dz(fd[2], dom, o.bcZType, AD::DIRZ, st);
        AD *dv[3] = {&dx, &dy, &dz};
        mc3d::MeshSpec m{};
        for (int d = 0; d < 3; ++d) {
            m.N[d] = dims[d];
            m.periodic[d] = bc[d] == MC3D_PERIODIC;
            m.bw7[d] = fd[d] == MC3D_PENTA10;
            for (int e = 0; e < 3; ++e) m.T[d][e] = trans && m.periodic[d] ? trans[3 * d + e] : 0.0;
        }
        mc3d::solve_for_jacobians(dv, m, x, y, z, out, st);
    });
}

int mc3d_solver_set_initial(mc3d_solver *s, const double *rho0, const double *U0, const double *V0,
                            const double *W0, const double *p0) {
    if (!s || !rho0 || !U0 || !V0 || !W0 || !p0) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (!s->metrics) return mc3d::fail(MC3D_EINVAL, "set the metrics before the initial condition");
    return guarded([&] { s->cs->setInitialConditions_host(rho0, U0, V0, W0, p0); s->initial = true; });
}

int mc3d_solver_load_restart(mc3d_solver *s) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    if (!s->metrics) return mc3d::fail(MC3D_EINVAL, "set the metrics before the initial state");
    if (!s->cs->opt.fromRestart) return mc3d::fail(MC3D_ECONFIG, "RESTART.FROMRESTART is not set");
    return guarded([&] {
        s->cs->load_restart_host(s->cs->opt.filename);
        s->initial = true;
    });
}

int mc3d_solver_write_dump(mc3d_solver *s, const char *path, const double *x, const double *y,
                           const double *z) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] {
        const double *mesh[3] = {x, y, z};
        s->cs->dumpSolution(path ? std::string(path) : std::string(), mesh);
    });
}

int mc3d_solver_set_output(mc3d_solver *s, const char *dir, const double *x, const double *y,
                           const double *z) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] {
        const double *mesh[3] = {x, y, z};
        s->cs->set_output(dir ? std::string(dir) : std::string(), mesh);
    });
}

int mc3d_solver_set_turbdata(mc3d_solver *s, const char *path, int every) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] { s->cs->set_turbdata(path ? std::string(path) : std::string(), every); });
}

int mc3d_dump_read_header(const char *path, int *timestep, double *time, int dims[3]) {
    if (!path) return mc3d::fail(MC3D_EINVAL, "null path");
    return guarded([&] {
        const mc3d::DumpHeader h = mc3d::read_dump_header(path);
        if (timestep) *timestep = h.timeStep;
        if (time) *time = h.time;
        if (dims) { dims[0] = h.Nx; dims[1] = h.Ny; dims[2] = h.Nz; }
    });
}

int mc3d_dump_read(const char *path, double *const fields[8]) {
    if (!path || !fields) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { mc3d::read_dump_fields(path, mc3d::read_dump_header(path), fields); });
}

int mc3d_dump_write(const char *path, int timestep, double time, const int dims[3],
                    const double *const fields[8]) {
    if (!path || !dims || !fields) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return mc3d::fail(MC3D_EINVAL, "dims must be positive");
    return guarded([&] {
        mc3d::DumpHeader h;
        h.timeStep = timestep; h.time = time; h.Nx = dims[0]; h.Ny = dims[1]; h.Nz = dims[2];
        mc3d::write_dump(path, h, fields);
    });
}

namespace {
int ready(const mc3d_solver *s) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    if (!s->initial) return mc3d::fail(MC3D_EINVAL, "initial condition not set");
    if (s->cs->hasSponge() && !s->sponge)
        return mc3d::fail(MC3D_ECONFIG, "SPONGE faces need the sponge strength: mc3d_solver_set_mesh or "
                                        "mc3d_solver_set_sponge");
    return MC3D_OK;
}
}  // namespace

int mc3d_solver_set_sponge(mc3d_solver *s, const double *sigma, const double *const avg[5]) {
    if (!s || !sigma) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (avg)
        for (int q = 0; q < 5; ++q)
            if (!avg[q]) return mc3d::fail(MC3D_EINVAL, "null sponge average");
    const int rc = guarded([&] { s->cs->setSponge(sigma, avg); });
    if (rc == MC3D_OK) s->sponge = true;
    return rc;
}

int mc3d_solver_step(mc3d_solver *s, int nsteps) {
    if (const int rc = ready(s)) return rc;
    if (nsteps < 0) return mc3d::fail(MC3D_EINVAL, "negative step count");
    return guarded([&] {
        for (int i = 0; i < nsteps; ++i) s->cs->rk->step();
    });
}

int mc3d_solver_run(mc3d_solver *s) {
    if (const int rc = ready(s)) return rc;
    return guarded([&] { s->cs->rk->executeSolverLoop(); });
}

int mc3d_solver_field_device(mc3d_solver *s, const char *name, double **dev_ptr) {
    if (!s || !name || !dev_ptr) return mc3d::fail(MC3D_EINVAL, "null argument");
    double *p = s->cs->field_by_name(name);
    if (!p) return mc3d::fail(MC3D_EINVAL, std::string("unknown field ") + name);
    *dev_ptr = p;
    return MC3D_OK;
}

int mc3d_solver_get_field(mc3d_solver *s, const char *name, double *host_out) {
    double *p = nullptr;
    int rc = mc3d_solver_field_device(s, name, &p);
    if (rc) return rc;
    if (!host_out) return mc3d::fail(MC3D_EINVAL, "null output");
    return guarded([&] {
        mc3d::hip_check(hipMemcpyAsync(host_out, p, sizeof(double) * s->cs->N, hipMemcpyDeviceToHost,
                                       s->stream), "field download");
        mc3d::hip_check(hipStreamSynchronize(s->stream), "field download");
    });
}

int mc3d_solver_add_image(mc3d_solver *s, int interval, int width, int height, const char *field,
                          const char *var_name, int plane, double fraction, int has_range, double vmin,
                          double vmax, int colormap) {
    if (!s || !field || !var_name) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] {
        mc3d::ImageOutput img;
        img.interval = interval;
        img.width = width;
        img.height = height;
        img.field = field;
        img.var_name = var_name;
        img.plane = plane;
        img.fraction = fraction;
        img.has_range = has_range != 0;
        img.vmin = vmin;
        img.vmax = vmax;
        img.colormap = colormap;
        s->cs->addImageOutput(img);
    });
}

int mc3d_solver_slice(mc3d_solver *s, const char *field, int dir, int index, double *host_out) {
    if (!s || !field || !host_out) return mc3d::fail(MC3D_EINVAL, "null argument");
    return guarded([&] { s->cs->slice(field, dir, index, host_out); });
}

int mc3d_image_write(const char *path, const double *plane, int n1, int n2, int width, int height,
                     int has_range, double vmin, double vmax, int colormap) {
    if (!path || !plane) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (n1 < 1 || n2 < 1 || width < 1 || height < 1) return mc3d::fail(MC3D_EINVAL, "sizes must be positive");
    if (colormap < mc3d::GREYSCALE || colormap > mc3d::BWR) return mc3d::fail(MC3D_EINVAL, "unknown colormap");
    mc3d::ImageOutput img;
    img.width = width;
    img.height = height;
    img.has_range = has_range != 0;
    img.vmin = vmin;
    img.vmax = vmax;
    img.colormap = colormap;
    const std::vector<unsigned char> rgb = mc3d::render_plane(plane, n1, n2, img);
    if (!mc3d::write_png(path, rgb.data(), width, height))
        return mc3d::fail(MC3D_EINVAL, std::string("cannot write ") + path);
    return MC3D_OK;
}

int mc3d_solver_scalars(mc3d_solver *s, double *time, double *dt, int *timestep) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] {
        mc3d::StepScalars h;
        mc3d::hip_check(hipMemcpyAsync(&h, s->cs->sc, sizeof(h), hipMemcpyDeviceToHost, s->stream), "scalar read");
        mc3d::hip_check(hipStreamSynchronize(s->stream), "scalar read");
        if (time) *time = h.time;
        if (dt) *dt = h.dt;
        if (timestep) *timestep = s->cs->timeStep;
    });
}

int mc3d_solver_stream(mc3d_solver *s, void **stream) {
    if (!s || !stream) return mc3d::fail(MC3D_EINVAL, "null argument");
    *stream = s->stream;
    return MC3D_OK;
}

int mc3d_solver_sync(mc3d_solver *s) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] { mc3d::hip_check(hipStreamSynchronize(s->stream), "synchronize"); });
}

int mc3d_solver_timing(mc3d_solver *s, int enable) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    return guarded([&] {
        s->cs->timer.reset();
        s->cs->timer.enable(enable != 0);
    });
}

int mc3d_solver_timing_report(mc3d_solver *s, char *buf, int buflen, int *needed) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    std::string j;
    int rc = guarded([&] {
        s->cs->timer.collect();
        j = "[";
        bool first = true;
        char tmp[512];
        for (const auto &kv : s->cs->timer.stats()) {
            std::snprintf(tmp, sizeof(tmp), "%s{\"kernel\": \"%s\", \"count\": %ld, \"ms\": %.6f, \"bytes\": %.0f}",
                          first ? "" : ", ", kv.first.c_str(), kv.second.count, kv.second.ms,
                          kv.second.bytes);
            j += tmp;
            first = false;
        }
        j += "]";
    });
    if (rc) return rc;
    if (needed) *needed = (int)j.size() + 1;
    if (!buf || buflen < (int)j.size() + 1) return mc3d::fail(MC3D_EINVAL, "report buffer too small");
    std::memcpy(buf, j.c_str(), j.size() + 1);
    return MC3D_OK;
}

int mc3d_solver_stats_weights(mc3d_solver *s, double *velocity, double *thermo) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    const mc3d::Stats *st = s->cs->stats.get();
    if (velocity) *velocity = st ? st->velocityStatsWeight : 0.0;
    if (thermo) *thermo = st ? st->thermoStatsWeight : 0.0;
    return MC3D_OK;
}

int mc3d_solver_set_closure(mc3d_solver *s, const mc3d_closure *c) {
    if (!s || !c) return mc3d::fail(MC3D_EINVAL, "null argument");
    auto *l = dynamic_cast<mc3d::LearnedSGS *>(s->cs->les.get());
    if (!l) return mc3d::fail(MC3D_ECONFIG, "LES.LESMODEL is not LEARNED");
    return guarded([&] { l->set_network(*mc3d_closure_net(c)); });
}

int mc3d_solver_sample_closure(mc3d_solver *s, float *X, float *T, long count, long offset, long stride) {
    if (!s || !X || !T) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (!s->initial) return mc3d::fail(MC3D_EINVAL, "set the initial condition first");
    if (count <= 0 || offset < 0 || stride <= 0 || offset + (count - 1) * stride >= s->cs->N)
        return mc3d::fail(MC3D_EINVAL, "sample points outside the field");
    return guarded([&] { s->cs->sampleClosure(X, T, count, offset, stride); });
}

int mc3d_solver_hit_stats(mc3d_solver *s, double stats[11]) {
    if (!s || !stats) return mc3d::fail(MC3D_EINVAL, "null argument");
    if (!s->initial) return mc3d::fail(MC3D_EINVAL, "set the initial condition first");
    return guarded([&] { s->cs->hitStats(stats); });
}

int mc3d_solver_eval_sgs(mc3d_solver *s) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    if (!s->initial) return mc3d::fail(MC3D_EINVAL, "set the initial condition first");
    return guarded([&] { s->cs->evalSGS(); });
}

int mc3d_solver_closure_stats(mc3d_solver *s, long *evaluations, long *outside_points, long *fallback_stages,
                              long *fallback_points) {
    if (!s) return mc3d::fail(MC3D_EINVAL, "null solver");
    auto *l = dynamic_cast<mc3d::LearnedSGS *>(s->cs->les.get());
    if (!l) return mc3d::fail(MC3D_ECONFIG, "LES.LESMODEL is not LEARNED");
    if (evaluations) *evaluations = l->evaluations;
    if (outside_points) *outside_points = l->outside_points;
    if (fallback_stages) *fallback_stages = l->fallback_stages;
    if (fallback_points) *fallback_points = l->fallback_points;
    return MC3D_OK;
}

void mc3d_solver_destroy(mc3d_solver *s) {
    if (!s) return;
    try {
        delete s->cs;
    } catch (...) {
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
