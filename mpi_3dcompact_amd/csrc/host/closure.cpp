// Learned SGS closure, host side (see closure.hpp).
#include "closure.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

namespace mc3d {

namespace {

constexpr char kMagic[8] = {'M', 'C', '3', 'D', 'M', 'L', 'P', '1'};

template <class T>
T *dev_alloc(size_t count, const char *what) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(T) * count), what);
    return static_cast<T *>(p);
}

}  // namespace

// He-style uniform initialisation, weights U(-s, s) with s = sqrt(3 / fan_in),
// biases zero (std::mt19937_64 with the given seed)
Closure::Closure(uint64_t seed) {
    params = dev_alloc<float>(kMlpParams, "closure parameters");
    m_ = dev_alloc<float>(kMlpParams, "closure moments");
    v_ = dev_alloc<float>(kMlpParams, "closure moments");
    grad_ = dev_alloc<float>(kMlpParams, "closure gradient");
    std::vector<float> h(kMlpParams, 0.0f);
    std::mt19937_64 rng(seed);
    const int fan[4] = {kMlpIn, kMlpHid, kMlpHid, kMlpHid};
    const int woff[4] = {kMlpOffW1, kMlpOffW2, kMlpOffW3, kMlpOffW4};
    const int wlen[4] = {kMlpHid * kMlpIn, kMlpHid * kMlpHid, kMlpHid * kMlpHid, kMlpOut * kMlpHid};
    for (int l = 0; l < 4; ++l) {
        std::uniform_real_distribution<double> u(-std::sqrt(3.0 / fan[l]), std::sqrt(3.0 / fan[l]));
        for (int q = 0; q < wlen[l]; ++q) h[woff[l] + q] = (float)u(rng);
    }
    set_params_host(h.data());
    for (int k = 0; k < kMlpIn; ++k) {
        sc.in_mean[k] = 0.0f; sc.in_std[k] = 1.0f;
        sc.lo[k] = -INFINITY; sc.hi[k] = INFINITY;
    }
    for (int k = 0; k < kMlpOut; ++k) { sc.out_mean[k] = 0.0f; sc.out_std[k] = 1.0f; }
}

Closure::~Closure() {
    (void)hipDeviceSynchronize();
    for (void *p : {(void *)params, (void *)m_, (void *)v_, (void *)grad_, (void *)loss_, (void *)partial_,
                    (void *)lossp_, (void *)idx_})
        if (p) (void)hipFree(p);
}

void Closure::set_params_host(const float *p) {
    hip_check(hipMemcpy(params, p, sizeof(float) * kMlpParams, hipMemcpyHostToDevice), "closure upload");
    reset_adam();
}

void Closure::get_params_host(float *p) const {
    hip_check(hipMemcpy(p, params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost), "closure download");
}

void Closure::reset_adam() {
    hip_check(hipMemset(m_, 0, sizeof(float) * kMlpParams), "closure moments");
    hip_check(hipMemset(v_, 0, sizeof(float) * kMlpParams), "closure moments");
    adam_t = 0;
}

void Closure::ensure(long batch, long n_idx, long nbatches) {
    const long nb = (batch + kMlpTile - 1) / kMlpTile;
    if (nb > part_blocks_) {
        if (partial_) (void)hipFree(partial_);
        if (lossp_) (void)hipFree(lossp_);
        partial_ = dev_alloc<float>((size_t)nb * kMlpParams, "closure partial gradients");
        lossp_ = dev_alloc<float>((size_t)nb, "closure partial losses");
        part_blocks_ = nb;
    }
    if (n_idx > idx_cap_) {
        if (idx_) (void)hipFree(idx_);
        idx_ = dev_alloc<int>((size_t)n_idx, "closure sample order");
        idx_cap_ = n_idx;
    }
    if (nbatches > loss_cap_) {
        if (loss_) (void)hipFree(loss_);
        loss_ = dev_alloc<float>((size_t)nbatches, "closure losses");
        loss_cap_ = nbatches;
    }
}

double Closure::gradient(const float *X, const float *T, long n, float *grad, hipStream_t st) {
    if (n <= 0) throw Error(MC3D_EINVAL, "closure gradient: no samples");
    ensure(n, 0, 1);
    const int nb = (int)((n + kMlpTile - 1) / kMlpTile);
    hip_check(launch_mlp_grad_partials(params, X, T, nullptr, n, n, partial_, lossp_, st), "closure gradient");
    hip_check(launch_mlp_reduce(partial_, lossp_, nb, n, grad, loss_, st), "closure gradient sum");
    float l = 0.0f;
    hip_check(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, st), "closure loss");
    hip_check(hipStreamSynchronize(st), "closure gradient");
    return l;
}

// Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8; bias corrections
// formed in FP64 and rounded to FP32
void Closure::adam(const float *grad, float lr, hipStream_t st) {
    ++adam_t;
    AdamArgs a{lr, 0.9f, 0.999f, 1e-8f, 0.0f, 0.0f};
    a.c1 = (float)(1.0 - std::pow((double)a.b1, adam_t));
    a.c2 = (float)(1.0 - std::pow((double)a.b2, adam_t));
    hip_check(launch_mlp_adam(params, m_, v_, grad, a, st), "closure adam");
}

void Closure::train(const float *X, const float *T, long n, int epochs, int batch, float lr, uint64_t seed,
                    double *losses, hipStream_t st) {
    if (n <= 0 || epochs <= 0 || batch <= 0) throw Error(MC3D_EINVAL, "closure training: empty problem");
    if (n > 0x7fffffffL) throw Error(MC3D_EINVAL, "closure training: more than 2^31 - 1 samples");
    const long nbatches = (n + batch - 1) / batch;
    ensure(batch, n, nbatches);
    std::vector<int> order((size_t)n);
    std::vector<float> bl((size_t)nbatches);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; ++e) {
        for (long q = 0; q < n; ++q) order[(size_t)q] = (int)q;
        for (long q = n - 1; q > 0; --q) {  // Fisher-Yates
            std::uniform_int_distribution<long> u(0, q);
            std::swap(order[(size_t)q], order[(size_t)u(rng)]);
        }
        hip_check(hipMemcpyAsync(idx_, order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st),
                  "closure sample order");
        for (long b = 0; b < nbatches; ++b) {
            const long nb_s = std::min<long>(batch, n - b * batch);
            const int nblk = (int)((nb_s + kMlpTile - 1) / kMlpTile);
            hip_check(launch_mlp_grad_partials(params, X, T, idx_ + b * batch, nb_s, nb_s, partial_, lossp_, st),
                      "closure gradient");
            hip_check(launch_mlp_reduce(partial_, lossp_, nblk, nb_s, grad_, loss_ + b, st), "closure gradient sum");
            adam(grad_, lr, st);
        }
        hip_check(hipMemcpyAsync(bl.data(), loss_, sizeof(float) * (size_t)nbatches, hipMemcpyDeviceToHost, st),
                  "closure losses");
        hip_check(hipStreamSynchronize(st), "closure epoch");
        double acc = 0.0;
        for (long b = 0; b < nbatches; ++b) acc += (double)bl[(size_t)b] * (double)std::min<long>(batch, n - b * batch);
        if (losses) losses[e] = acc / (double)n;
    }
}

void Closure::infer(const float *X, float *Y, long n, hipStream_t st) const {
    hip_check(launch_mlp_infer(params, X, Y, n, st), "closure inference");
}

// file: 8-byte magic, int32 parameter count, FP32 parameters, MlpScaling (FP32)
void Closure::save(const std::string &path) const {
    std::vector<float> p(kMlpParams);
    get_params_host(p.data());
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(MC3D_EINVAL, "cannot write " + path);
    const int32_t np = kMlpParams;
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&np, 4, 1, f) == 1 &&
              std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size() &&
              std::fwrite(&sc, sizeof(sc), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw Error(MC3D_EINVAL, "short write to " + path);
}

void Closure::load(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(MC3D_EINVAL, "cannot open " + path);
    char magic[8];
    int32_t np = 0;
    std::vector<float> p(kMlpParams);
    MlpScaling s{};
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                    std::fread(&np, 4, 1, f) == 1 && np == kMlpParams &&
                    std::fread(p.data(), sizeof(float), p.size(), f) == p.size() &&
                    std::fread(&s, sizeof(s), 1, f) == 1;
    std::fclose(f);
    if (!ok) throw Error(MC3D_EINVAL, path + " is not a closure network file");
    set_params_host(p.data());
    sc = s;
    scaled = true;
}

// ---------------------------------------------------------------------------
// The network's inputs are averaged as the DSM it replaces averages L_ij M_ij
// and M_ij M_ij (LESAVERAGING, src/DSMSGS.hpp:245-300 with src/AbstractSGS.cpp:13-164):
// line means, the global mean, or the test filter for LOCAL.  A network is to
// be used with the averaging its training pairs were sampled under.
LearnedSGS::LearnedSGS(CurvilinearCSolver *c) {
    cs = c;
    fallback = c->opt.closureFallback;
    // CLOSUREFEATURES POINT: the network reads the features at each point
    // whatever LESAVERAGING says (which still sets the fallback DSM's averaging)
    musgsAvgType = c->opt.closurePointFeatures ? Options::NONE : c->opt.musgsAvgType;
    mu_sgs = c->alloc_field();
    k_sgs = c->alloc_field();
    if (musgsAvgType != Options::NONE) {
        for (auto &f : feat_) f = c->alloc_field();
        smag_ = c->alloc_field();
        delta2_ = c->alloc_field();
    }
    if (musgsAvgType == Options::XI1_AVG || musgsAvgType == Options::XI2_AVG || musgsAvgType == Options::XI3_AVG) {
        line_dir_ = c->axis_of[musgsAvgType == Options::XI1_AVG ? 0 : musgsAvgType == Options::XI2_AVG ? 1 : 2];
        const Domain &d = c->dom;
        nlines_ = (long)d.gNx * d.gNy * d.gNz / (line_dir_ == 0 ? d.gNx : line_dir_ == 1 ? d.gNy : d.gNz);
        for (auto &f : lines_) f = dev_alloc<double>((size_t)nlines_, "closure line means");
        line_C_ = dev_alloc<float>((size_t)nlines_, "closure line coefficients");
        line_D_ = dev_alloc<float>((size_t)nlines_, "closure line coefficients");
        line_flag_ = dev_alloc<unsigned char>((size_t)nlines_, "closure line flags");
    }
    if (musgsAvgType == Options::LOCAL) {
        for (auto &f : ftmp_) f = c->alloc_field();
        const int bc3[3] = {c->opt.bcXType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                            c->opt.bcYType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET,
                            c->opt.bcZType == Options::PERIODIC_SOLVE ? MC3D_PERIODIC : MC3D_DIRICHLET};
        std::unique_ptr<AbstractFilter> *fv[3] = {&filtX, &filtY, &filtZ};
        for (int d = 0; d < 3; ++d) {
            const auto dir = (AbstractDerivatives::Direct)c->axis_of[d];
            if (c->opt.testFilterType == 1) fv[d]->reset(new CommN5PadeTestFilter(c->dom, bc3, dir, c->st));
            else fv[d]->reset(new CommN5ExpTestFilter(c->dom, bc3, dir, c->st));
            static_cast<LineFilter *>(fv[d]->get())->timer = &c->timer;
        }
    }
    params_ = dev_alloc<float>(kMlpParams, "closure parameters");
    flag_ = dev_alloc<unsigned char>((size_t)c->N, "closure flags");
    blockcount_ = dev_alloc<int>((size_t)closure_blocks(c->N), "closure counts");
    count_ = dev_alloc<int>(1, "closure count");
    if (fallback) dsm_.reset(new DSMSGS(c, true));
    if (!c->opt.closureFile.empty()) {
        Closure net(0);
        net.load(c->opt.closureFile);
        set_network(net);
    }
}

LearnedSGS::~LearnedSGS() {
    (void)hipStreamSynchronize(cs->st);
    for (void *p : {(void *)params_, (void *)flag_, (void *)blockcount_, (void *)count_, (void *)line_C_,
                    (void *)line_D_, (void *)line_flag_})
        if (p) (void)hipFree(p);
    for (double *p : lines_)
        if (p) (void)hipFree(p);
}

void LearnedSGS::set_network(const Closure &c) {
    if (!c.scaled) throw Error(MC3D_EINVAL, "the closure network has no standardisation set");
    hip_check(hipMemcpyAsync(params_, c.params, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice, cs->st),
              "closure copy");
    hip_check(hipStreamSynchronize(cs->st), "closure copy");
    sc_ = c.sc;
    ready_ = true;
}

void LearnedSGS::getSGSViscosity(const double *const dU[3], const double *const dV[3], const double *const dW[3],
                                 const double *rho) {
    if (!ready_) throw Error(MC3D_ECONFIG, "LESMODEL LEARNED needs a closure network (CLOSUREFILE or "
                                           "mc3d_solver_set_closure)");
    ClosureArgs a{};
    for (int d = 0; d < 3; ++d) { a.dU[d] = dU[d]; a.dV[d] = dV[d]; a.dW[d] = dW[d]; }
    a.m = cs->msh;
    a.rho = rho;
    a.params = params_;
    a.sc = sc_;
    a.vol = cs->dom.dx * cs->dom.dy * cs->dom.dz;
    a.cp = cs->ig.cp;
    a.mu_sgs = mu_sgs;
    a.k_sgs = k_sgs;
    a.flag = flag_;
    a.blockcount = blockcount_;
    int count_blocks = 0;  // workgroup counts the flagged-point total is summed over
    if (feat_[0]) {
        FeatureArgs fa{};
        for (int d = 0; d < 3; ++d) { fa.dU[d] = dU[d]; fa.dV[d] = dV[d]; fa.dW[d] = dW[d]; }
        fa.m = cs->msh;
        fa.vol = a.vol;
        for (int k = 0; k < kMlpIn; ++k) fa.feat[k] = feat_[k];
        fa.smag = smag_;
        fa.delta2 = delta2_;
        // read: 9 derivatives, 10 metrics; written: 7 features, |S|, Delta^2
        cs->timed("k_closure_features", 28.0,
                  [&] { hip_check(launch_closure_features(fa, cs->N, cs->st), "closure features"); });
        if (line_dir_ >= 0) {
            // line means of the features, the network once per line, then every point
            const Domain &d = cs->dom;
            const int len = line_dir_ == 0 ? d.gNx : line_dir_ == 1 ? d.gNy : d.gNz;
            cs->timed("k_xi_average_fields", 7.0, [&] {
                hip_check(launch_xi_average_fields(feat_, nullptr, kMlpIn, line_dir_, d.gNx, d.gNy, d.gNz, cs->st,
                                                   lines_),
                          "closure feature average");
            });
            LineClosureArgs la{};
            for (int k = 0; k < kMlpIn; ++k) la.feat[k] = lines_[k];
            la.params = params_;
            la.sc = sc_;
            la.nlines = nlines_;
            la.line_len = len;
            la.C = line_C_;
            la.D = line_D_;
            la.lflag = line_flag_;
            la.blockcount = blockcount_;
            cs->timed("k_line_closure", 0.0, [&] { hip_check(launch_line_closure(la, cs->st), "line closure"); });
            ClosureApplyArgs ap{};
            ap.rho = rho;
            ap.smag = smag_;
            ap.delta2 = delta2_;
            ap.C = line_C_;
            ap.D = line_D_;
            ap.lflag = line_flag_;
            ap.cp = cs->ig.cp;
            ap.dir = line_dir_;
            ap.Nx = d.gNx;
            ap.Ny = d.gNy;
            ap.mu_sgs = mu_sgs;
            ap.k_sgs = k_sgs;
            ap.flag = flag_;
            // read: rho, |S|, Delta^2; written: mu_sgs, k_sgs
            cs->timed("k_closure_apply", 5.0,
                      [&] { hip_check(launch_closure_apply(ap, cs->N, cs->st), "closure apply"); });
            count_blocks = line_closure_blocks(nlines_);
        } else {
            if (musgsAvgType == Options::LOCAL) {
                // doAveraging's test filter (Y, then Z in place, then X) of
                // every feature, one launch per direction
                const double *in[kMlpIn], *tc[kMlpIn];
                for (int k = 0; k < kMlpIn; ++k) in[k] = feat_[k], tc[k] = ftmp_[k];
                filtY->filterFields(kMlpIn, in, ftmp_);
                filtZ->filterFields(kMlpIn, tc, ftmp_);
                filtX->filterFields(kMlpIn, tc, feat_);
            } else {
                for (int k = 0; k < kMlpIn; ++k) doAveraging(feat_[k], feat_[k]);
            }
            for (int k = 0; k < kMlpIn; ++k) a.feat[k] = feat_[k];
            a.smag = smag_;
            a.delta2 = delta2_;
        }
    }
    if (count_blocks == 0) {
        // read: 9 derivatives, 10 metrics, rho (or 7 averaged features, |S|, Delta^2, rho);
        // written: mu_sgs, k_sgs (+ one byte of flags)
        cs->timed("k_learned_sgs", feat_[0] ? 12.0 : 22.0,
                  [&] { hip_check(launch_learned_sgs(a, cs->N, cs->st), "learned closure"); });
        count_blocks = closure_blocks(cs->N);
    }
    ++evaluations;
    if (!fallback) return;
    hip_check(launch_count_sum(blockcount_, count_blocks, count_, cs->st), "closure count");
    int n_out = 0;
    hip_check(hipMemcpyAsync(&n_out, count_, sizeof(int), hipMemcpyDeviceToHost, cs->st), "closure count");
    hip_check(hipStreamSynchronize(cs->st), "closure count");
    outside_points += n_out;
    if (n_out == 0 || (double)n_out <= cs->opt.closureFallbackFraction * (double)cs->N) return;
    fallback_points += n_out;
    ++fallback_stages;
    // the DSM on this stage (without tau_kk, which the closure does not model),
    // merged where the inputs left the trained range
    const Cons &s = cs->state();
    dsm_->calcMuSGSTaukk(dU, dV, dW, s.q[0], s.q[1], s.q[2], s.q[3]);
    dsm_->calcKSGS(cs->dT, s.q[0], s.q[1], s.q[2], s.q[3], cs->pr.T);
    FallbackArgs f{flag_, dsm_->mu_sgs, dsm_->k_sgs, mu_sgs, k_sgs};
    cs->timed("k_closure_fallback", 4.0, [&] { hip_check(launch_closure_fallback(f, cs->N, cs->st), "fallback"); });
}

}  // namespace mc3d
