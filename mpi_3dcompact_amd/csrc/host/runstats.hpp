// Running statistics of the solution (src/Stats.hpp:9-263): time-weighted
// means, RMS values and Reynolds stresses of U, V, W and of rho, p, T kept on
// the device, updated every STATS.INTERVAL steps and written beside every
// SolutionDump.
#pragma once

#include <string>

#include "options.hpp"

namespace mc3d {

class CurvilinearCSolver;

class Stats {
  public:
    explicit Stats(CurvilinearCSolver *cs);
    ~Stats();
    Stats(const Stats &) = delete;
    Stats &operator=(const Stats &) = delete;

    bool velocityStats = false, thermoStats = false;
    double velocityStatsWeight = 0.0, thermoStatsWeight = 0.0;
    Options::StatsAvgType statsAvgType = Options::NONE;
    double prev_time = 0.0;
    // device fields in file order: UAVG VAVG WAVG URMS VRMS WRMS UVREY UWREY
    // VWREY, and RHOAVG RHORMS PAVG PRMS TAVG TRMS (null when off)
    double *vel[9] = {}, *thermo[6] = {};

    // src/Stats.cpp:121-246, at the solver time `time`
    void updateStatsFields(double time);
    // src/Stats.cpp:3-119: VelocityStats.<timeStep> and ThermoStats.<timeStep>
    // in dir (empty: the working directory)
    void dumpStatsFields(const std::string &dir, int timeStep);
    // src/Stats.cpp:248-345
    void doAveraging(double *phi);
    // device field of a statistics name (UAVG .. TRMS), null if unknown or off
    double *field(const std::string &name) const;

  private:
    CurvilinearCSolver *cs_;
    void *mem_ = nullptr;
    void load(const std::string &path, double &weight, double *const *f, int nf);
};

}  // namespace mc3d
