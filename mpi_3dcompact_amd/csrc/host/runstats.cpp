// Running statistics (runstats.hpp).
#include "runstats.hpp"

#include <cstdio>
#include <fstream>
#include <vector>

#include "../runstats.h"
#include "solver.hpp"

namespace mc3d {

namespace {
const char *const kVelNames[9] = {"UAVG", "VAVG", "WAVG", "URMS", "VRMS", "WRMS", "UVREY", "UWREY", "VWREY"};
const char *const kThermoNames[6] = {"RHOAVG", "RHORMS", "PAVG", "PRMS", "TAVG", "TRMS"};
}  // namespace

// src/Stats.hpp:42-250: fields start at zero with weight zero, or from the
// statistics files named by the options (weight, then the fields)
Stats::Stats(CurvilinearCSolver *cs) : cs_(cs) {
    const Options &o = cs->opt;
    velocityStats = o.velocityStats;
    thermoStats = o.thermoStats;
    statsAvgType = o.statsAvgType;
    prev_time = o.time;
    const long N = cs->N;
    const int nf = (velocityStats ? 9 : 0) + (thermoStats ? 6 : 0);
    if (nf == 0) return;
    hip_check(hipMalloc(&mem_, sizeof(double) * (size_t)N * nf), "statistics allocation");
    hip_check(hipMemsetAsync(mem_, 0, sizeof(double) * (size_t)N * nf, cs->st), "statistics clear");
    double *p = static_cast<double *>(mem_);
    if (velocityStats)
        for (auto &f : vel) { f = p; p += N; }
    if (thermoStats)
        for (auto &f : thermo) { f = p; p += N; }
    if (velocityStats && o.velocityStatsFromRestart) load(o.velocityStatsFilename, velocityStatsWeight, vel, 9);
    if (thermoStats && o.thermoStatsFromRestart) load(o.thermoStatsFilename, thermoStatsWeight, thermo, 6);
    hip_check(hipStreamSynchronize(cs->st), "statistics set-up");
    if (velocityStats) std::printf(" > Initialized Velocity Stats, Weight = %g\n", velocityStatsWeight);
    if (thermoStats) std::printf(" > Initialized Thermo Stats, Weight = %g\n", thermoStatsWeight);
}

Stats::~Stats() {
    if (mem_) {
        (void)hipStreamSynchronize(cs_->st);
        (void)hipFree(mem_);
    }
}

void Stats::load(const std::string &path, double &weight, double *const *f, int nf) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(MC3D_EINVAL, "cannot open statistics file " + path);
    const long N = cs_->N;
    std::vector<double> buf((size_t)N);
    in.read(reinterpret_cast<char *>(&weight), sizeof(double));
    for (int k = 0; k < nf; ++k) {
        in.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)(sizeof(double) * N));
        if (!in) throw Error(MC3D_EINVAL, "statistics file " + path + " is shorter than its grid");
        hip_check(hipMemcpy(f[k], buf.data(), sizeof(double) * N, hipMemcpyHostToDevice), "statistics upload");
    }
}

void Stats::doAveraging(double *phi) {
    const Domain &d = cs_->dom;
    if (statsAvgType == Options::XI1_AVG)
        hip_check(launch_xi_average(phi, phi, 0, d.gNx, d.gNy, d.gNz, cs_->st), "statistics average");
    else if (statsAvgType == Options::XI3_AVG)
        hip_check(launch_xi_average(phi, phi, 2, d.gNx, d.gNy, d.gNz, cs_->st), "statistics average");
    // XI2_AVG averages a scratch array of the reference and leaves phi unchanged
    // (src/Stats.cpp:274-297); NONE, LOCAL and GLOBAL do nothing
}

void Stats::updateStatsFields(double time) {
    const double dt = time - prev_time;
    const long N = cs_->N;
    if (velocityStats && dt != 0.0) {
        const double oldWeight = velocityStatsWeight, totNewWeight = oldWeight + dt;
        StatsArgs a{};
        a.a = cs_->pr.U; a.b = cs_->pr.V; a.c = cs_->pr.W;
        for (int k = 0; k < 9; ++k) a.f[k] = vel[k];
        a.oldW = oldWeight; a.dt = dt; a.totW = totNewWeight;
        hip_check(launch_stats_velocity(a, N, cs_->st), "velocity statistics");
        for (int k : {3, 4, 5, 6, 7, 8, 0, 1, 2}) doAveraging(vel[k]);
        hip_check(launch_stats_velocity_finish(a, N, cs_->st), "velocity statistics");
        std::printf(" > Updated velocity stats, current stats weight = %g\n", totNewWeight);
        velocityStatsWeight = totNewWeight;
    }
    if (thermoStats && dt != 0.0) {
        const double oldWeight = thermoStatsWeight, totNewWeight = oldWeight + dt;
        StatsArgs a{};
        // rho2: the last stage's register before the filter (src/Stats.hpp:58)
        a.a = cs_->r2.q[0]; a.b = cs_->pr.p; a.c = cs_->pr.T;
        for (int k = 0; k < 6; ++k) a.f[k] = thermo[k];
        a.oldW = oldWeight; a.dt = dt; a.totW = totNewWeight;
        hip_check(launch_stats_thermo(a, N, cs_->st), "thermo statistics");
        for (int k : {1, 3, 5, 0, 2, 4}) doAveraging(thermo[k]);
        hip_check(launch_stats_thermo_finish(a, N, cs_->st), "thermo statistics");
        std::printf(" > Updated thermo stats, current stats weight = %g\n", totNewWeight);
        thermoStatsWeight = totNewWeight;
    }
    prev_time = time;
}

void Stats::dumpStatsFields(const std::string &dir, int timeStep) {
    const long N = cs_->N;
    auto write = [&](const char *stem, double weight, double *const *f, int nf) {
        const std::string name = (dir.empty() ? std::string() : dir + "/") + stem + std::to_string(timeStep);
        std::ofstream out(name, std::ios::binary | std::ios::trunc);
        if (!out) throw Error(MC3D_EINVAL, "cannot write " + name);
        out.write(reinterpret_cast<const char *>(&weight), sizeof(double));
        std::vector<double> buf((size_t)N);
        for (int k = 0; k < nf; ++k) {
            hip_check(hipMemcpyAsync(buf.data(), f[k], sizeof(double) * N, hipMemcpyDeviceToHost, cs_->st),
                      "statistics read-back");
            hip_check(hipStreamSynchronize(cs_->st), "statistics read-back");
            out.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(sizeof(double) * N));
        }
        if (!out) throw Error(MC3D_EINVAL, "write failed: " + name);
    };
    if (velocityStats) {
        std::printf("\n > ========================\n >  DUMPING VELOCITY STATS \n > ========================\n");
        write("VelocityStats.", velocityStatsWeight, vel, 9);
    }
    if (thermoStats) {
        std::printf("\n > ======================\n >  DUMPING THERMO STATS \n > ======================\n");
        write("ThermoStats.", thermoStatsWeight, thermo, 6);
    }
}

double *Stats::field(const std::string &name) const {
    for (int k = 0; k < 9; ++k)
        if (name == kVelNames[k]) return vel[k];
    for (int k = 0; k < 6; ++k)
        if (name == kThermoNames[k]) return thermo[k];
    return nullptr;
}

}  // namespace mc3d
