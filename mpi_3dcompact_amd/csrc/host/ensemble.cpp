// Ensembles of periodic boxes: buckets of stacked members, graph-captured
// steps and the line-operator shape autotuner (ensemble.hpp).
#include "ensemble.hpp"

#include <array>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "../capi_common.h"

namespace mc3d {

namespace {

const char *fd_name(Options::FDType t) { return t == Options::CD2 ? "CD2" : t == Options::PADE6 ? "PADE6" : "PENTA10"; }

int fd_scheme(Options::FDType t) {
    return t == Options::CD2 ? MC3D_CD2 : t == Options::PADE6 ? MC3D_PADE6 : MC3D_PENTA10;
}

// one cache line: "<key> = <mm> <family> <lines> <ms> <candidates>"
bool parse_shape(const std::string &v, TunedShape &t) {
    std::istringstream is(v);
    return static_cast<bool>(is >> t.shape.mm >> t.shape.family >> t.shape.lines >> t.ms >> t.candidates);
}

std::string format_shape(const TunedShape &t) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%d %d %d %.6f %d", t.shape.mm, t.shape.family, t.shape.lines, t.ms,
                  t.candidates);
    return buf;
}

std::string json_escape(const std::string &s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

}  // namespace

Ensemble::Ensemble(const Options &opt, int members, const int *dims, const double *mu) : opt_(opt) {
    if (members < 1) throw Error(MC3D_EINVAL, "an ensemble needs at least one member");
    if (!dims) throw Error(MC3D_EINVAL, "null member sizes");
    if (opt.fromRestart || opt.onlyGridFromRestart)
        throw Error(MC3D_ECONFIG, "ensemble members start from their initial conditions, not a restart file");
    for (int m = 0; m < members; ++m) {
        for (int k = 0; k < 3; ++k)
            if (dims[3 * m + k] < 8) throw Error(MC3D_EINVAL, "every member dimension must be at least 8");
        if (mu && !(mu[m] >= 0.0)) throw Error(MC3D_EINVAL, "MU_REF of a member must be non-negative");
    }
    // buckets by box size, in order of first appearance
    std::map<std::array<int, 3>, int> index;
    where_.resize((size_t)members);
    for (int m = 0; m < members; ++m) {
        const std::array<int, 3> d{dims[3 * m], dims[3 * m + 1], dims[3 * m + 2]};
        auto it = index.find(d);
        if (it == index.end()) {
            it = index.emplace(d, (int)buckets_.size()).first;
            buckets_.emplace_back();
            Bucket &b = buckets_.back();
            b.dims[0] = d[0]; b.dims[1] = d[1]; b.dims[2] = d[2];
        }
        Bucket &b = buckets_[(size_t)it->second];
        where_[(size_t)m] = {it->second, (int)b.members.size()};
        b.members.push_back(m);
    }
    have_metrics_.assign((size_t)members, 0);
    have_initial_.assign((size_t)members, 0);
    hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "stream creation");
    hip_check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "event creation");
    for (Bucket &b : buckets_) {
        Options o = opt_;
        o.Nx = b.dims[0]; o.Ny = b.dims[1]; o.Nz = b.dims[2];
        std::vector<double> mus;
        for (int m : b.members) mus.push_back(mu ? mu[m] : opt_.mu_ref);
        hip_check(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking), "stream creation");
        hip_check(hipEventCreateWithFlags(&b.done, hipEventDisableTiming), "event creation");
        // one member alone takes its MU_REF through the options
        if (b.members.size() == 1) o.mu_ref = mus[0];
        b.cs.reset(new CurvilinearCSolver(o, b.st, (int)b.members.size(), mus.data()));
    }
}

Ensemble::~Ensemble() {
    if (st_) (void)hipStreamSynchronize(st_);
    for (auto &kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    for (Bucket &b : buckets_) {
        if (b.st) (void)hipStreamSynchronize(b.st);
        b.cs.reset();
        if (b.done) (void)hipEventDestroy(b.done);
        if (b.st) (void)hipStreamDestroy(b.st);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (st_) (void)hipStreamDestroy(st_);
}

void Ensemble::setMetrics(int m, const double *const metrics[10]) {
    if (m < 0 || m >= members()) throw Error(MC3D_EINVAL, "member index out of range");
    for (int k = 0; k < 10; ++k)
        if (!metrics[k]) throw Error(MC3D_EINVAL, "null metric array");
    sync();
    const auto w = where_[(size_t)m];
    buckets_[(size_t)w.first].cs->setMemberMetrics(w.second, metrics);
    have_metrics_[(size_t)m] = 1;
}

void Ensemble::setInitial(int m, const double *rho, const double *U, const double *V, const double *W,
                          const double *p) {
    if (m < 0 || m >= members()) throw Error(MC3D_EINVAL, "member index out of range");
    if (!rho || !U || !V || !W || !p) throw Error(MC3D_EINVAL, "null initial field");
    if (timeStep() != opt_.timeStep) throw Error(MC3D_EINVAL, "initial conditions are set before the first step");
    sync();
    const auto w = where_[(size_t)m];
    buckets_[(size_t)w.first].cs->setMemberInitial(w.second, rho, U, V, W, p);
    have_initial_[(size_t)m] = 1;
}

void Ensemble::checkReady() const {
    for (int m = 0; m < (int)where_.size(); ++m) {
        if (!have_metrics_[(size_t)m]) throw Error(MC3D_EINVAL, "member " + std::to_string(m) + ": metrics not set");
        if (!have_initial_[(size_t)m])
            throw Error(MC3D_EINVAL, "member " + std::to_string(m) + ": initial condition not set");
    }
}

int Ensemble::timeStep() const { return buckets_.front().cs->timeStep; }

void Ensemble::advanceAll() {
    hip_check(hipEventRecord(fork_, st_), "ensemble fork");
    for (Bucket &b : buckets_) {
        hip_check(hipStreamWaitEvent(b.st, fork_, 0), "ensemble fork");
        b.cs->rk->advance();
        hip_check(hipEventRecord(b.done, b.st), "ensemble join");
        hip_check(hipStreamWaitEvent(st_, b.done, 0), "ensemble join");
    }
}

void Ensemble::step(int nsteps) {
    if (nsteps < 0) throw Error(MC3D_EINVAL, "negative step count");
    checkReady();
    for (int s = 0; s < nsteps; ++s) {
        const int key = buckets_.front().cs->stepKey();
        if (!graphs_on_ || key < 0) {
            advanceAll();
            continue;
        }
        auto it = graphs_.find(key);
        if (it == graphs_.end()) {
            // capture runs the host side of the step (its bookkeeping included) and
            // records its launches without executing them
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "graph capture");
            try {
                advanceAll();
            } catch (...) {
                hipGraph_t junk = nullptr;
                (void)hipStreamEndCapture(st_, &junk);
                if (junk) (void)hipGraphDestroy(junk);
                throw;
            }
            hip_check(hipStreamEndCapture(st_, &g), "graph capture");
            hipGraphExec_t x = nullptr;
            const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e, "graph instantiation");
            it = graphs_.emplace(key, x).first;
        } else {
            for (Bucket &b : buckets_) b.cs->skipAdvance();
        }
        hip_check(hipGraphLaunch(it->second, st_), "graph launch");
        ++graph_launches_;
    }
}

void Ensemble::useGraphs(bool on) { graphs_on_ = on; }

void Ensemble::sync() {
    hip_check(hipStreamSynchronize(st_), "ensemble synchronize");
    for (Bucket &b : buckets_) hip_check(hipStreamSynchronize(b.st), "ensemble synchronize");
}

void Ensemble::getField(int m, const std::string &name, double *host) {
    if (m < 0 || m >= members()) throw Error(MC3D_EINVAL, "member index out of range");
    if (!host) throw Error(MC3D_EINVAL, "null output");
    sync();
    const auto w = where_[(size_t)m];
    Bucket &b = buckets_[(size_t)w.first];
    const double *p = b.cs->memberField(w.second, name);
    if (!p) throw Error(MC3D_EINVAL, "unknown or absent field: " + name);
    hip_check(hipMemcpyAsync(host, p, sizeof(double) * b.cs->dom.gN, hipMemcpyDeviceToHost, b.st), "field read");
    hip_check(hipStreamSynchronize(b.st), "field read");
}

StepScalars Ensemble::scalars(int m) {
    if (m < 0 || m >= members()) throw Error(MC3D_EINVAL, "member index out of range");
    sync();
    const auto w = where_[(size_t)m];
    return buckets_[(size_t)w.first].cs->memberScalars(w.second);
}

void Ensemble::dropGraphs() {
    sync();
    for (auto &kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

// ---------------------------------------------------------------------------
// Autotuner.  For every bucket, direction and operator (derivative, filter)
// each launch shape the line operators accept is built and timed with HIP
// events on the bucket's stream over the operator's use in a step: a
// derivative of 4 fields (velocity and temperature gradients) plus a
// derivative of 5 fields added into the divergence; a filter of the 5
// conserved fields.  Outputs go to work fields the step overwrites anyway.
void Ensemble::tune(const std::string &cache_path) {
    dropGraphs();
    std::map<std::string, std::string> cache;
    if (!cache_path.empty()) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            const size_t eq = line.find(" = ");
            if (eq != std::string::npos) cache[line.substr(0, eq)] = line.substr(eq + 3);
        }
    }
    for (Bucket &b : buckets_) tuneBucket(b, cache);
    if (!cache_path.empty()) {
        std::ofstream f(cache_path, std::ios::trunc);
        if (!f) throw Error(MC3D_EINVAL, "cannot write the tuning cache " + cache_path);
        f << "# line-operator launch shapes: <operator box members dir> = <mm> <family> <lines> <ms> <candidates>\n";
        for (auto &kv : cache) f << kv.first << " = " << kv.second << "\n";
    }
}

void Ensemble::tuneBucket(Bucket &b, std::map<std::string, std::string> &cache) {
    CurvilinearCSolver &cs = *b.cs;
    hip_check(hipStreamSynchronize(b.st), "tuning");
    const Options::FDType ft[3] = {opt_.xFDType, opt_.yFDType, opt_.zFDType};
    const double dd[3] = {cs.dom.dx, cs.dom.dy, cs.dom.dz};
    const int bc3[3] = {MC3D_PERIODIC, MC3D_PERIODIC, MC3D_PERIODIC};
    const int fkind = opt_.filterType == Options::COMPACT8 ? MC3D_COMPACT8 : MC3D_COMPACT10;
    const int nm = cs.members();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event creation");
    hip_check(hipEventCreate(&e1), "event creation");
    LineTune dsh[3], fsh[3];
    b.tuned.clear();
    for (int cls = 0; cls < 2; ++cls)
        for (int d = 0; d < 3; ++d) {
            char key[256];
            if (cls == 0)
                std::snprintf(key, sizeof(key), "deriv %s periodic %dx%dx%d members %d dir %d", fd_name(ft[d]),
                              b.dims[0], b.dims[1], b.dims[2], nm, d);
            else
                std::snprintf(key, sizeof(key), "filter %s alpha %.6g periodic %dx%dx%d members %d dir %d",
                              fkind == MC3D_COMPACT8 ? "COMPACT8" : "COMPACT10", opt_.alphaF, b.dims[0],
                              b.dims[1], b.dims[2], nm, d);
            TunedShape best;
            best.key = key;
            auto hit = cache.find(key);
            if (hit != cache.end() && parse_shape(hit->second, best)) {
                best.cached = true;
            } else {
                best.ms = -1.0;
                best.candidates = 0;
                const int kind = cls == 0 ? fd_scheme(ft[d]) : fkind;
                auto time_op = [&](const mc3d_lineop *op) {
                    auto once = [&] {
                        if (cls == 0) {
                            const double *in4[4] = {cs.pr.U, cs.pr.V, cs.pr.W, cs.pr.T};
                            double *out4[4] = {cs.dU[d], cs.dV[d], cs.dW[d], cs.dT[d]};
                            if (lineop_launch_accumulate(op, in4, out4, 4, 0, b.st) != MC3D_OK)
                                throw Error(MC3D_EHIP, "tuning launch failed");
                            const double *in5[5] = {cs.pre[5 * d], cs.pre[5 * d + 1], cs.pre[5 * d + 2],
                                                    cs.pre[5 * d + 3], cs.pre[5 * d + 4]};
                            if (lineop_launch_accumulate(op, in5, cs.div, 5, 1, b.st) != MC3D_OK)
                                throw Error(MC3D_EHIP, "tuning launch failed");
                        } else {
                            const double *in5[5] = {cs.r1.q[0], cs.r1.q[1], cs.r1.q[2], cs.r1.q[3], cs.r1.q[4]};
                            double *out5[5] = {cs.tmp[0], cs.tmp[1], cs.tmp[2], cs.tmp[3], cs.tmp[4]};
                            if (lineop_launch_accumulate(op, in5, out5, 5, 0, b.st) != MC3D_OK)
                                throw Error(MC3D_EHIP, "tuning launch failed");
                        }
                    };
                    once();  // warm-up
                    const int reps = 3;
                    hip_check(hipEventRecord(e0, b.st), "tuning");
                    for (int r = 0; r < reps; ++r) once();
                    hip_check(hipEventRecord(e1, b.st), "tuning");
                    hip_check(hipEventSynchronize(e1), "tuning");
                    float ms = 0.f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "tuning");
                    return (double)ms / reps;
                };
                // candidates: the untuned default first (the plan's own
                // chunking, which for line lengths such as 40, 72 or 200 is an
                // even 8..28-row split that 16- and 32-row chunks do not
                // give), then 16- and 32-row chunks for the fast kernels and
                // for the LDS-tiled kernel with 1..256 lines per workgroup
                std::vector<LineTune> cand(1);
                for (int mm : {16, 32})
                    for (int family : {1, 2})
                        for (int lines = family == 1 ? 0 : 1; lines <= (family == 1 ? 0 : 256);
                             lines = family == 1 ? 1 : 2 * lines) {
                            LineTune t;
                            t.mm = mm;
                            t.family = family;
                            t.lines = lines;
                            cand.push_back(t);
                        }
                for (const LineTune &t : cand) {
                    mc3d_lineop *op = nullptr;
                    if (lineop_create(cls == 0 ? kDeriv : kFilter, kind, d, bc3, b.dims, dd[d], opt_.alphaF, nm,
                                      t, &op) != MC3D_OK)
                        continue;
                    double ms = 0.0;
                    try {
                        ms = time_op(op);
                    } catch (...) {
                        mc3d_lineop_destroy(op);
                        throw;
                    }
                    mc3d_lineop_destroy(op);
                    ++best.candidates;
                    best.trials.emplace_back(t, ms);
                    if (best.ms < 0.0 || ms < best.ms) {
                        best.ms = ms;
                        best.shape = t;
                    }
                }
                if (best.candidates == 0) throw Error(MC3D_EINVAL, std::string("no launch shape for ") + key);
                cache[key] = format_shape(best);
            }
            (cls == 0 ? dsh : fsh)[d] = best.shape;
            b.tuned.push_back(best);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cs.setLineShapes(dsh, fsh);
}

std::string Ensemble::info() const {
    std::ostringstream o;
    o << "{\"members\": " << where_.size() << ", \"graphs\": " << graphs_.size()
      << ", \"graph_launches\": " << graph_launches_ << ", \"buckets\": [";
    for (size_t i = 0; i < buckets_.size(); ++i) {
        const Bucket &b = buckets_[i];
        const CurvilinearCSolver &cs = *b.cs;
        o << (i ? ", " : "") << "{\"dims\": [" << b.dims[0] << ", " << b.dims[1] << ", " << b.dims[2]
          << "], \"members\": [";
        for (size_t k = 0; k < b.members.size(); ++k) o << (k ? ", " : "") << b.members[k];
        o << "], \"deriv\": [";
        const AbstractDerivatives *dv[3] = {cs.derivX.get(), cs.derivY.get(), cs.derivZ.get()};
        const AbstractFilter *fv[3] = {cs.filtX.get(), cs.filtY.get(), cs.filtZ.get()};
        for (int d = 0; d < 3; ++d)
            o << (d ? ", " : "") << "\""
              << json_escape(lineop_kernel_name(static_cast<const LineDerivative *>(dv[d])->op())) << "\"";
        o << "], \"filter\": [";
        for (int d = 0; d < 3; ++d)
            o << (d ? ", " : "") << "\""
              << json_escape(lineop_kernel_name(static_cast<const LineFilter *>(fv[d])->op())) << "\"";
        o << "], \"tuned\": [";
        for (size_t k = 0; k < b.tuned.size(); ++k) {
            const TunedShape &t = b.tuned[k];
            o << (k ? ", " : "") << "{\"key\": \"" << json_escape(t.key) << "\", \"mm\": " << t.shape.mm
              << ", \"family\": " << t.shape.family << ", \"lines\": " << t.shape.lines << ", \"ms\": " << t.ms
              << ", \"candidates\": " << t.candidates << ", \"cached\": " << (t.cached ? "true" : "false")
              << ", \"trials\": [";
            for (size_t r = 0; r < t.trials.size(); ++r)
                o << (r ? ", " : "") << "[" << t.trials[r].first.mm << ", " << t.trials[r].first.family << ", "
                  << t.trials[r].first.lines << ", " << t.trials[r].second << "]";
            o << "]}";
        }
        o << "]}";
    }
    o << "]}";
    return o.str();
}

}  // namespace mc3d
