// Host C++ side of the hot path: the reference's plugin interfaces
// (AbstractDerivatives, AbstractFilter, AbstractRK, AbstractSGS) re-declared
// over device pointers, their HIP-backed implementations, and the
// CurvilinearCSolver orchestrator that calls them in the reference's order.
// Every operator runs on the solver's HIP stream; nothing here touches field
// data on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../lineop_api.h"
#include "../pointwise.h"
#include "mc3d.h"
#include "images.hpp"
#include "options.hpp"
#include "runstats.hpp"
#include "timer.hpp"

namespace mc3d {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what);

// Computational domain with unit length per direction (src/Domain.hpp:31-89,
// src/MPI_3DCompact.cpp:132): spacing 1/N periodic, 1/(N-1) Dirichlet.
struct Domain {
    int gNx, gNy, gNz;
    long gN;          // points of one member
    double dx, dy, dz;
    int members = 1;  // boxes of this size stacked along z in every field
    explicit Domain(const Options &o);
};

// ---------------------------------------------------------------------------
// Derivatives, src/AbstractDerivatives.hpp:13-53
// ---------------------------------------------------------------------------
class AbstractDerivatives {
  public:
    enum Direct { DIRX, DIRY, DIRZ };
    enum BW { BW3, BW5, BW7 };
    int Nx, Ny, Nz, N;
    double dd;
    Direct currentDir;
    Options::BCType bcType;
    BW lhsBandwidth, rhsBandwidth;

    virtual ~AbstractDerivatives() = default;
    // d(dataIn)/d(xi_dir) into dataOut (device pointers)
    virtual void calc1stDerivField(const double *dataIn, double *dataOut) = 0;
    // transformed periodic boundaries, rhs bandwidth 5 and 7
    virtual void calc1stDerivField_TPB(const double *dataIn, double *dataOut, const double *Nm2,
                                       const double *Nm1, const double *Np1, const double *Np2) = 0;
    virtual void calc1stDerivField_TPB(const double *dataIn, double *dataOut, const double *Nm3,
                                       const double *Nm2, const double *Nm1, const double *Np1,
                                       const double *Np2, const double *Np3) = 0;
    // one-sided wall value (host values, as the reference evaluates it)
    virtual double calcNeumann(const double *f) = 0;
    // batched form: n fields in one launch; accumulate adds into dataOut
    virtual void calc1stDerivFields(int n, const double *const *in, double *const *out,
                                    bool accumulate) = 0;
};

class LineDerivative : public AbstractDerivatives {
  public:
    LineDerivative(int scheme, const Domain &dom, Options::BCType bc, Direct dir, hipStream_t st,
                   const LineTune &shape = LineTune{});
    ~LineDerivative() override;
    void calc1stDerivField(const double *in, double *out) override;
    void calc1stDerivField_TPB(const double *in, double *out, const double *Nm2, const double *Nm1,
                               const double *Np1, const double *Np2) override;
    void calc1stDerivField_TPB(const double *in, double *out, const double *Nm3, const double *Nm2,
                               const double *Nm1, const double *Np1, const double *Np2,
                               const double *Np3) override;
    double calcNeumann(const double *f) override;
    void calc1stDerivFields(int n, const double *const *in, double *const *out, bool accumulate) override;
    KernelTimer *timer = nullptr;
    const mc3d_lineop *op() const { return op_; }

  protected:
    int scheme_;
    mc3d_lineop *op_ = nullptr;
    hipStream_t st_;
};

// src/Pade6.hpp:9-259, src/Penta10.hpp:9-197, src/CD2.hpp:9-83
struct Pade6 : LineDerivative {
    Pade6(const Domain &d, Options::BCType bc, Direct dir, hipStream_t st, const LineTune &t = LineTune{})
        : LineDerivative(MC3D_PADE6, d, bc, dir, st, t) {}
};
struct Penta10 : LineDerivative {
    Penta10(const Domain &d, Options::BCType bc, Direct dir, hipStream_t st, const LineTune &t = LineTune{})
        : LineDerivative(MC3D_PENTA10, d, bc, dir, st, t) {}
};
struct CD2 : LineDerivative {
    CD2(const Domain &d, Options::BCType bc, Direct dir, hipStream_t st, const LineTune &t = LineTune{})
        : LineDerivative(MC3D_CD2, d, bc, dir, st, t) {}
};

// ---------------------------------------------------------------------------
// Filters, src/AbstractFilter.hpp:13-35
// ---------------------------------------------------------------------------
class AbstractFilter {
  public:
    int Nx, Ny, Nz, N;
    AbstractDerivatives::Direct currentDir;
    Options::BCType bcType;
    virtual ~AbstractFilter() = default;
    virtual void filterField(const double *dataIn, double *dataOut) = 0;
    // batched form (in may equal out)
    virtual void filterFields(int n, const double *const *in, double *const *out) = 0;
};

class LineFilter : public AbstractFilter {
  public:
    LineFilter(int kind, double alphaF, const Domain &dom, const int bc3[3],
               AbstractDerivatives::Direct dir, hipStream_t st, const LineTune &shape = LineTune{});
    ~LineFilter() override;
    void filterField(const double *in, double *out) override;
    void filterFields(int n, const double *const *in, double *const *out) override;
    KernelTimer *timer = nullptr;
    const mc3d_lineop *op() const { return op_; }

  protected:
    mc3d_lineop *op_ = nullptr;
    hipStream_t st_;
};

// src/Compact8Filter.hpp:9-139, src/Compact10Filter.hpp:9-308,
// src/CommN5PadeTestFilter.hpp, src/CommN5ExpTestFilter.hpp
struct Compact8Filter : LineFilter {
    Compact8Filter(double af, const Domain &d, const int bc3[3], AbstractDerivatives::Direct dir, hipStream_t st,
                   const LineTune &t = LineTune{})
        : LineFilter(MC3D_COMPACT8, af, d, bc3, dir, st, t) {}
};
struct Compact10Filter : LineFilter {
    Compact10Filter(double af, const Domain &d, const int bc3[3], AbstractDerivatives::Direct dir, hipStream_t st,
                    const LineTune &t = LineTune{})
        : LineFilter(MC3D_COMPACT10, af, d, bc3, dir, st, t) {}
};
struct CommN5PadeTestFilter : LineFilter {
    CommN5PadeTestFilter(const Domain &d, const int bc3[3], AbstractDerivatives::Direct dir, hipStream_t st)
        : LineFilter(MC3D_COMMN5PADE, 0.0, d, bc3, dir, st) {}
};
struct CommN5ExpTestFilter : LineFilter {
    CommN5ExpTestFilter(const Domain &d, const int bc3[3], AbstractDerivatives::Direct dir, hipStream_t st)
        : LineFilter(MC3D_COMMN5EXP, 0.0, d, bc3, dir, st) {}
};

class CurvilinearCSolver;

// ---------------------------------------------------------------------------
// SGS models, src/AbstractSGS.hpp:9-42
// ---------------------------------------------------------------------------
class AbstractSGS {
  public:
    CurvilinearCSolver *cs = nullptr;
    Options::StatsAvgType musgsAvgType = Options::NONE;
    std::unique_ptr<AbstractFilter> filtX, filtY, filtZ;
    double *mu_sgs = nullptr, *taukk = nullptr, *k_sgs = nullptr;
    bool coeffRangeFlag = false;
    double *C = nullptr, *CI = nullptr, *Prt = nullptr;
    virtual ~AbstractSGS() = default;
    // gradients are passed as the computational-space derivatives dU[3], dV[3],
    // dW[3]; the Cartesian tensor is formed in the kernels that use it.
    virtual void getSGSViscosity(const double *const dU[3], const double *const dV[3],
                                 const double *const dW[3], const double *rho) = 0;
    virtual void calcMuSGSTaukk(const double *const dU[3], const double *const dV[3],
                                const double *const dW[3], const double *rho, const double *rhoU,
                                const double *rhoV, const double *rhoW) = 0;
    virtual void calcKSGS(const double *const dT[3], const double *rho, const double *rhoU,
                          const double *rhoV, const double *rhoW, const double *T) = 0;
    // src/AbstractSGS.cpp:13-164
    void doAveraging(const double *phi, double *phiF);
    bool isDSM = false;
};

class VremanSGS : public AbstractSGS {
  public:
    explicit VremanSGS(CurvilinearCSolver *cs);
    void getSGSViscosity(const double *const dU[3], const double *const dV[3],
                         const double *const dW[3], const double *rho) override;
    void calcMuSGSTaukk(const double *const *, const double *const *, const double *const *,
                        const double *, const double *, const double *, const double *) override {}
    void calcKSGS(const double *const *, const double *, const double *, const double *,
                  const double *, const double *) override {}
};

// Dynamic Smagorinsky model, src/DSMSGS.hpp:13-451.  The test filter is
// CommN5PadeTestFilter (DSMTESTFILTER 1) or CommN5ExpTestFilter (2); every
// filterQuantity / doAveraging of the reference (Y, then Z, then X) is issued
// as batched line-operator launches over all the fields it filters.
class DSMSGS : public AbstractSGS {
  public:
    // fallback_only: the learned closure's fallback (no tau_kk, no coefficient output)
    explicit DSMSGS(CurvilinearCSolver *cs, bool fallback_only = false);
    void getSGSViscosity(const double *const *, const double *const *, const double *const *,
                         const double *) override {}
    void calcMuSGSTaukk(const double *const dU[3], const double *const dV[3],
                        const double *const dW[3], const double *rho, const double *rhoU,
                        const double *rhoV, const double *rhoW) override;
    void calcKSGS(const double *const dT[3], const double *rho, const double *rhoU,
                  const double *rhoV, const double *rhoW, const double *T) override;

    bool useTaukk = true;
    double *S[6] = {}, *Smag = nullptr, *rho_hat = nullptr, *Smag_hat = nullptr;
    double *rhoU_hat = nullptr, *rhoV_hat = nullptr, *rhoW_hat = nullptr;
    double *work[19] = {};  // work[1..18] as the reference numbers them
    double *gT[3] = {};     // Cartesian temperature gradient (filtered in calcKSGS)

  private:
    double *ftmp_[16] = {};
    // filterQuantity of n fields: Y into temporaries, Z in place, X into out
    void filterQuantity(int n, const double *const *in, double *const *out);
    // doAveraging of n fields (batched for LOCAL)
    void average(int n, const double *const *in, double *const *out);
};

// ---------------------------------------------------------------------------
// RK integrators, src/AbstractRK.hpp:6-17
// ---------------------------------------------------------------------------
class AbstractRK {
  public:
    CurvilinearCSolver *cs = nullptr;
    virtual ~AbstractRK() = default;
    // runs full steps until the solver's end condition (maxTimeStep / maxTime)
    virtual void executeSolverLoop();
    // one full time step: advance(), then the solver's checks, dumps and end test
    void step();
    // the launches of one full step (CFL and all stages) with their host
    // bookkeeping, and nothing that waits for the device
    void advance();
    // register update of the current stage (fused with the k2 assembly)
    virtual void updateConservedData() = 0;
    virtual int stages() const = 0;
    // which of the scheme's step forms comes next (LSLDDRK4 alternates two)
    virtual int variant() const { return 0; }
    // host bookkeeping at the end of a step
    virtual void finishStep() {}
};

class TVDRK3 : public AbstractRK {  // src/TVDRK3.hpp
  public:
    explicit TVDRK3(CurvilinearCSolver *c) { cs = c; }
    void updateConservedData() override;
    int stages() const override { return 3; }
};
class RK4 : public AbstractRK {  // src/RK4.hpp
  public:
    explicit RK4(CurvilinearCSolver *c) { cs = c; }
    void updateConservedData() override;
    int stages() const override { return 4; }
};
class KenRK4 : public AbstractRK {  // src/KenRK4.hpp
  public:
    explicit KenRK4(CurvilinearCSolver *c) { cs = c; }
    void updateConservedData() override;
    int stages() const override { return 5; }
};
class LSLDDRK4 : public AbstractRK {  // src/LSLDDRK4.hpp: alternating 5- and 6-stage steps
  public:
    explicit LSLDDRK4(CurvilinearCSolver *c) { cs = c; }
    void updateConservedData() override;
    int stages() const override { return rkDualStepStep == 1 ? 5 : 6; }
    int variant() const override { return rkDualStepStep - 1; }
    void finishStep() override { rkDualStepStep = rkDualStepStep == 1 ? 2 : 1; }
    int rkDualStepStep = 1;
};

// ---------------------------------------------------------------------------
// Solver, src/CurvilinearCSolver.hpp:25-379
// ---------------------------------------------------------------------------
class CurvilinearCSolver {
  public:
    // members > 1: that many boxes of the options' size stacked along z in
    // every field, each with its own time step and, when mu_members is given,
    // its own MU_REF (periodic boxes, no LES model or Vreman with NONE / LOCAL
    // averaging)
    CurvilinearCSolver(const Options &opt, hipStream_t st, int members = 1, const double *mu_members = nullptr);
    // The SGS model of the options alone, over caller-owned fields (the
    // mc3d_sgs handle): no flow derivatives, filters, RK or field slab beyond
    // the model's own work fields.  The fields are in `layout` (MC3D_XPENCIL,
    // YPENCIL or ZPENCIL): the domain takes the layout's axis order, axis_of
    // maps each physical direction to its axis, and the Cartesian gradients
    // are given (no metric terms: msh holds the caller's J alone).
    struct SgsOnly {
        int layout;
        const double *J;
    };
    CurvilinearCSolver(const Options &opt, hipStream_t st, SgsOnly t);
    ~CurvilinearCSolver();
    // axis of the fields along physical direction d (x, y, z)
    int axis_of[3] = {0, 1, 2};

    Options opt;
    Domain dom;
    Gas ig;
    hipStream_t st;
    long N;
    int timeStep, filterTimeStep = 0, rkStep = 0;
    bool rkLast = false, endFlag = false, firstStep = true;
    double Pr_t = 0.9;

    std::unique_ptr<AbstractDerivatives> derivX, derivY, derivZ;
    std::unique_ptr<AbstractFilter> filtX, filtY, filtZ;
    std::unique_ptr<AbstractSGS> les;
    std::unique_ptr<AbstractRK> rk;
    std::unique_ptr<Stats> stats;  // STATS.VELOCITYSTATS / THERMOSTATS (src/CurvilinearCSolver.hpp:224-229)

    // device fields (one slab)
    Metrics msh{};
    double *Jm[10] = {};
    Cons r1{}, rk_{}, k2{}, r2{};
    Prim pr{};
    double *dU[3] = {}, *dV[3] = {}, *dW[3] = {}, *dT[3] = {};
    double *pre[15] = {}, *div[5] = {};
    double *tmp[10] = {};
    double *cfl_partial = nullptr;
    StepScalars *sc = nullptr;
    KernelTimer timer;
    // times one pointwise launch under its kernel name with the given
    // number of full fields read + written
    template <class F> void timed(const char *key, double fields, F &&launch) {
        int t = timer.begin(key, fields * 8.0 * (double)N);
        launch();
        timer.end(t);
    }

    double *alloc_field();
    void set_metrics_host(const double *const m[10]);
    // metrics from mesh coordinates (device x, y, z), src/AbstractSingleBlockMesh.cpp:23-1309
    void solveForJacobians(const double *x, const double *y, const double *z);
    void set_mesh_host(const double *x, const double *y, const double *z);
    void periodicTranslations(double T[3][3]) const;
    double *msh_rw(int k) { return Jm[k]; }
    // FROMRESTART: conserved fields of a SolutionDump file converted to the
    // primitive initial state (src/MPI_3DCompact.cpp:214-271), then set up
    void load_restart_host(const std::string &path);
    // SolutionDump of the current rho1..rhoE1 with the given mesh
    // (src/CurvilinearCSolver_Core.cpp:1333-1409); empty path: the
    // reference's name in the output directory
    void dumpSolution(const std::string &path, const double *const mesh[3]);
    // dumps every DUMPSTEP steps and at the end, as the reference's postStep
    // does, into dir (mesh arrays copied; null entries are written as zeros)
    void set_output(const std::string &dir, const double *const mesh[3]);
    // PngWriter image outputs (CurvilinearCSolver::addImageOutput,
    // src/CurvilinearCSolver_Core.cpp:1438-1440), written at the first step
    // and every interval steps after the dump (:1442-1492)
    void addImageOutput(const ImageOutput &img);
    void writeImages();
    // index plane `index` normal to dir of a field into host memory, as
    // (n2, n1) with a1 fastest (the image planes' axes)
    void slice(const std::string &name, int dir, int index, double *host_out);
    void setInitialConditions_host(const double *rho0, const double *U0, const double *V0,
                                   const double *W0, const double *p0);

    // stage methods, in the reference's call order
    void preStep();               // calcDtFromCFL
    void preSubStep();            // preStepDerivatives
    void gradientsAndSGS();       //   its derivative and SGS part
    void flux();                  //   its flux and divergence part
    void solveEqnSet() {}         // k2 assembly is fused into the RK stage kernel
    void postSubStep() {}         // periodic boundaries: no boundary fix-ups
    void updateData();            // filter (last stage / substeps) + EOS
    void postStep();              // checks, dumps, end conditions
    void checkSolution();         // field ranges every CHECKSTEP steps
    void getRange(const double *phi, const char *name);
    void fieldRange(const double *phi, double &lo, double &hi);
    // the metric identities I_k = d(J_1k)/dxi + d(J_2k)/deta + d(J_3k)/dzeta of
    // the current metrics with the solver's derivatives, as (min, max) pairs
    // (src/AbstractSingleBlockMesh.cpp:1336-1393); print: the reference's log lines
    void metricIdentity(double range[6], bool print);
    void stage_update(int op, double a, double b, double c);
    void filterConservedData();
    void updateNonConservedData();
    void run_stage(AbstractRK *r, int s, int S);

    const Cons &state() const { return rkStep == 1 ? r1 : rk_; }
    double *field_by_name(const std::string &name);

    // stacked members
    int members() const { return dom.members; }
    void setMemberInitial(int b, const double *rho0, const double *U0, const double *V0, const double *W0,
                          const double *p0);
    void setMemberMetrics(int b, const double *const m[10]);
    // member b's part of a full field (null for names that are not full fields)
    double *memberField(int b, const std::string &name);
    StepScalars memberScalars(int b);
    // rebuild the derivative and filter operators with the given launch shapes
    void setLineShapes(const LineTune deriv[3], const LineTune filt[3]);
    // launch sequence of the next step: steps with equal keys issue the same
    // launches with the same arguments (-1: the first step, which differs)
    int stepKey() const;
    // the host bookkeeping of advance() without its launches (for a replayed step)
    void skipAdvance();

    // boundary handling (src/CurvilinearCSolver_BC.cpp, src/SpongeBC.cpp)
    void setupBoundaries();
    void wallInitialConditions();   // :63-322
    void preStepBCHandling();       // :372-700
    RkBc rkBoundaryArgs() const;    // :702-842 and the sponge source, applied in the stage kernel
    void initSpongeAverages();      // :324-345 (from the current state unless given)
    // sponge strength sigma: from host mesh coordinates (SpongeBC::initRectSpongeBC /
    // initCylSpongeBC) or given; averages optional (null: the current state)
    void setSpongeFromMesh(const double *x, const double *y, const double *z);
    void setSponge(const double *sigma_host, const double *const avg_host[5]);
    bool hasSponge() const { return sponge_; }
    double *sigmaField() const { return sigma_; }
    double *spongeAvg(int q) const { return spgAvg_[q]; }
    // SpongeAvgDump.<timeStep> in dir: the five sponge averages (SpongeBC::dumpSpongeAvg,
    // src/SpongeBC.cpp:363-420)
    void dumpSpongeAvg(const std::string &dir);
    // learned closure: training pairs (features, DSM coefficients) of the
    // current state; the DSM runs on it and the fields its tau_kk update
    // changes are restored afterwards
    void sampleClosure(float *X, float *T, long count, long offset, long stride);
    // derivatives and SGS fields of the current state without advancing it
    // (models that leave the thermodynamic state unchanged)
    void evalSGS();
    // decaying-HIT statistics of the current state, the turbdata.out columns
    // of tests/hit/Samtaney_D4/MPI_3DCompact.cpp:545-556: time, tau,
    // Re_lambda, M_t, epsilon, u', Kolmogorov scale, rho', dil', omega',
    // <rho u.u>
    void hitStats(double out[11]);
    // the HIT driver's fullStepTemporalHook: append the hitStats line to
    // `path` every `every` steps (empty path: off)
    void set_turbdata(const std::string &path, int every);

  private:
    std::string turb_path_;
    int turb_every_ = 25;
    void initGas();
    std::vector<double> mu_members_;  // per-member MU_REF when members > 1
    std::chrono::steady_clock::time_point last_check_ = std::chrono::steady_clock::now();
    bool dumps_ = false;
    std::string out_dir_;
    std::vector<ImageOutput> images_;
    double *slice_buf_ = nullptr;
    size_t slice_cap_ = 0;
    int images_step_ = -1;  // the step whose images were last written
    std::vector<double> mesh_[3];
    bool eos_done_ = false;  // the last stage kernel also updated the non-conserved data
    // boundaries
    signed char role_[6] = {0, 0, 0, 0, 0, 0};
    bool walls_ = false, sponge_ = false, sigma_set_ = false, avg_given_ = false;
    double *cont_[3] = {nullptr, nullptr, nullptr};
    double *sigma_ = nullptr, *spgAvg_[5] = {};
    void *slab_ = nullptr;
    size_t slab_bytes_ = 0, slab_used_ = 0;
    std::vector<void *> extra_;
};

}  // namespace mc3d
