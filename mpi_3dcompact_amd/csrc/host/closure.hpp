// Learned SGS closure, host side: the network with its training state, and
// the AbstractSGS that evaluates it inside the solver step with the dynamic
// Smagorinsky model (src/DSMSGS.hpp) as the fallback on out-of-range inputs.
// No counterpart in the reference; it sits beside VremanSGS and DSMSGS behind
// the same AbstractSGS interface (src/AbstractSGS.hpp:9-42).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../closure.h"
#include "solver.hpp"

namespace mc3d {

class Closure {
  public:
    explicit Closure(uint64_t seed);
    ~Closure();
    Closure(const Closure &) = delete;
    Closure &operator=(const Closure &) = delete;

    float *params = nullptr;  // device FP32 [kMlpParams]
    MlpScaling sc{};
    bool scaled = false;
    int adam_t = 0;

    void set_params_host(const float *p);
    void get_params_host(float *p) const;
    void reset_adam();
    // full-batch gradient over X[n][7], T[n][2] (standardised, device) into
    // grad (device, kMlpParams); returns the loss
    double gradient(const float *X, const float *T, long n, float *grad, hipStream_t st);
    // one Adam step with a given gradient (device)
    void adam(const float *grad, float lr, hipStream_t st);
    // epochs of shuffled minibatch Adam; losses[e] = mean minibatch loss of epoch e
    void train(const float *X, const float *T, long n, int epochs, int batch, float lr, uint64_t seed,
               double *losses, hipStream_t st);
    void infer(const float *X, float *Y, long n, hipStream_t st) const;
    void save(const std::string &path) const;
    void load(const std::string &path);

  private:
    float *m_ = nullptr, *v_ = nullptr, *grad_ = nullptr, *loss_ = nullptr;
    float *partial_ = nullptr, *lossp_ = nullptr;
    int *idx_ = nullptr;
    long idx_cap_ = 0, part_blocks_ = 0, loss_cap_ = 0;
    void ensure(long batch, long n_idx, long nbatches);
};

class LearnedSGS : public AbstractSGS {
  public:
    explicit LearnedSGS(CurvilinearCSolver *cs);
    ~LearnedSGS() override;
    void getSGSViscosity(const double *const dU[3], const double *const dV[3], const double *const dW[3],
                         const double *rho) override;
    void calcMuSGSTaukk(const double *const *, const double *const *, const double *const *, const double *,
                        const double *, const double *, const double *) override {}
    void calcKSGS(const double *const *, const double *, const double *, const double *, const double *,
                  const double *) override {}
    // copies the network (parameters and scaling) into the solver
    void set_network(const Closure &c);
    bool has_network() const { return ready_; }
    bool fallback = true;
    // points outside the trusted range (all stages), and the stages and points
    // that fell back to the DSM (more than CLOSUREFALLBACKFRACTION of the points out)
    long outside_points = 0, fallback_points = 0, fallback_stages = 0, evaluations = 0;

  private:
    float *params_ = nullptr;
    MlpScaling sc_{};
    bool ready_ = false;
    unsigned char *flag_ = nullptr;
    int *blockcount_ = nullptr, *count_ = nullptr;
    std::unique_ptr<DSMSGS> dsm_;
    // averaged raw features (LESAVERAGING other than NONE), with |S| and Delta^2 per point
    double *feat_[kMlpIn] = {};
    double *smag_ = nullptr, *delta2_ = nullptr;
    // LOCAL: temporaries of the batched test filter of the features
    double *ftmp_[kMlpIn] = {};
    // XI1 / XI2 / XI3: the network per line (line means, coefficients, flags)
    int line_dir_ = -1;
    long nlines_ = 0;
    double *lines_[kMlpIn] = {};
    float *line_C_ = nullptr, *line_D_ = nullptr;
    unsigned char *line_flag_ = nullptr;
};

}  // namespace mc3d

// the network behind a C ABI closure handle (capi_closure.cpp)
struct mc3d_closure;
const mc3d::Closure *mc3d_closure_net(const mc3d_closure *c);
