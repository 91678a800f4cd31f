// Per-kernel HIP event timing (see timer.hpp).
#include "timer.hpp"

#include "solver.hpp"

namespace mc3d {

KernelTimer::~KernelTimer() {
    for (auto &r : recs_) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t KernelTimer::get_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "event creation");
    return e;
}

int KernelTimer::begin(const std::string &key, double bytes) {
    if (!on_) return -1;
    Rec r{key, bytes, get_event(), get_event()};
    hip_check(hipEventRecord(r.a, st_), "event record");
    recs_.push_back(r);
    return (int)recs_.size() - 1;
}

void KernelTimer::end(int token) {
    if (token < 0) return;
    hip_check(hipEventRecord(recs_[token].b, st_), "event record");
}

void KernelTimer::collect() {
    if (recs_.empty()) return;
    hip_check(hipStreamSynchronize(st_), "timer synchronize");
    for (auto &r : recs_) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, r.a, r.b), "event elapsed");
        Stat &s = stats_[r.key];
        s.count += 1;
        s.ms += ms;
        s.bytes += r.bytes;
        pool_.push_back(r.a);
        pool_.push_back(r.b);
    }
    recs_.clear();
}

}  // namespace mc3d
