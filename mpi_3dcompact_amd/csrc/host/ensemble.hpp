// Ensembles of independent periodic boxes of uneven sizes advanced together.
//
// Members of equal box size form a bucket: one CurvilinearCSolver whose fields
// hold the bucket's members stacked along z, so that every launch of its step
// covers all of them (x- and y-lines and pointwise work see one taller box;
// z-lines and the CFL reduction are per member; each member keeps its own dt,
// time and MU_REF).  Buckets run on streams of their own, forked from and
// joined into the ensemble stream; from the second step on, one step of the
// whole ensemble is captured into a HIP graph per launch sequence (the
// filter's direction order cycles through three) and replayed.
//
// The line-operator launch shapes (rows per chunk, kernel family, lines per
// workgroup) can be chosen per bucket and direction by measurement
// (tune()), with the results kept in a text cache keyed by operator, box size
// and member count.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "options.hpp"
#include "solver.hpp"

namespace mc3d {

struct TunedShape {
    std::string key;   // operator and box (the cache key)
    LineTune shape;
    double ms = 0.0;   // measured time of the usage pattern with this shape (0: default, not measured)
    int candidates = 0;
    bool cached = false;
    std::vector<std::pair<LineTune, double>> trials;  // every candidate measured (empty when cached)
};

class Ensemble {
  public:
    // dims: 3 per member; mu: MU_REF per member (null: the options' value)
    Ensemble(const Options &opt, int members, const int *dims, const double *mu);
    ~Ensemble();

    int members() const { return (int)where_.size(); }
    void setMetrics(int m, const double *const metrics[10]);
    void setInitial(int m, const double *rho, const double *U, const double *V, const double *W, const double *p);
    // nsteps full steps of every member (graph replays after the first step when enabled)
    void step(int nsteps);
    void useGraphs(bool on);
    void getField(int m, const std::string &name, double *host);
    StepScalars scalars(int m);
    int timeStep() const;
    void sync();
    // choose the line-operator shapes of every bucket by measurement (cache_path:
    // a text file read first and updated, or empty)
    void tune(const std::string &cache_path);
    // JSON description: buckets, members, kernels, tuning results
    std::string info() const;
    int graphLaunches() const { return graph_launches_; }
    hipStream_t stream() const { return st_; }
    int graphsCaptured() const { return (int)graphs_.size(); }

  private:
    struct Bucket {
        int dims[3];
        std::vector<int> members;  // ensemble indices, in slot order
        hipStream_t st = nullptr;
        hipEvent_t done = nullptr;
        std::unique_ptr<CurvilinearCSolver> cs;
        std::vector<TunedShape> tuned;  // deriv x,y,z then filter x,y,z once tuned
    };
    Options opt_;
    std::vector<Bucket> buckets_;
    std::vector<std::pair<int, int>> where_;  // member -> (bucket, slot)
    std::vector<char> have_metrics_, have_initial_;
    hipStream_t st_ = nullptr;
    hipEvent_t fork_ = nullptr;
    bool graphs_on_ = true;
    std::map<int, hipGraphExec_t> graphs_;
    int graph_launches_ = 0;

    void checkReady() const;
    void advanceAll();  // fork, every bucket's advance(), join (on st_)
    void dropGraphs();
    void tuneBucket(Bucket &b, std::map<std::string, std::string> &cache);
};

}  // namespace mc3d
