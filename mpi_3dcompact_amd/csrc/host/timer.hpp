// Per-kernel timing of the solver step with HIP events on the solver stream
// (the counterpart of the reference's SOLVER.USETIMING phase timers,
// src/CurvilinearCSolver_Core.cpp:6,182-185,...).  Each launch is recorded
// under its kernel's symbol name together with the algorithmic bytes it moves
// (distinct full fields read + written, 8 bytes per point each).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace mc3d {

class KernelTimer {
  public:
    struct Stat {
        long count = 0;
        double ms = 0.0;
        double bytes = 0.0;
    };
    explicit KernelTimer(hipStream_t st) : st_(st) {}
    ~KernelTimer();
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }
    // returns a token to pass to end(); -1 when disabled
    int begin(const std::string &key, double bytes);
    void end(int token);
    // synchronises, folds the recorded events into totals, clears the records
    void collect();
    void reset() { collect(); stats_.clear(); }
    const std::map<std::string, Stat> &stats() const { return stats_; }

  private:
    struct Rec {
        std::string key;
        double bytes;
        hipEvent_t a, b;
    };
    hipStream_t st_;
    bool on_ = false;
    std::vector<Rec> recs_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, Stat> stats_;
    hipEvent_t get_event();
};

}  // namespace mc3d
