// Options file (INI) of the solver: the reference's sections, keys, defaults,
// enum spellings and validation (src/Options.hpp:133-426, src/Options.cpp).
// Parsing errors are returned as text instead of aborting the process.
#pragma once

#include <map>
#include <string>

namespace mc3d {

struct Options {
    enum TimeSteppingType { CONST_DT, CONST_CFL };
    enum BCType { PERIODIC_SOLVE, DIRICHLET_SOLVE };
    enum BCKind { PERIODIC, INTERNALLY_PERIODIC, SPONGE, ADIABATIC_WALL, CONST_T_WALL,
                  MOVING_ADIABATIC_WALL, INLET };
    enum SpongeKind { RECTILINEAR, CYLINDRICAL };
    enum FDType { CD2, PADE6, PENTA10 };
    enum FilterType { COMPACT8, COMPACT10 };
    enum LESModel { NOMODEL, VREMAN, DSM, LEARNED };
    enum StatsAvgType { NONE, XI1_AVG, XI2_AVG, XI3_AVG, LOCAL, GLOBAL };
    enum RKType { TVDRK3, RK4, KENRK4, LSLDDRK4 };

    // [TIMESTEPPING]
    TimeSteppingType TSType = CONST_CFL;
    double CFL = 0.5, dt = -1.0, maxTime = 1000;
    int maxTimeStep = 10000, filterStep = 1, checkStep = 1, dumpStep = 1000;
    bool subStepFiltering = false;
    // [BC]
    BCType bcXType = PERIODIC_SOLVE, bcYType = PERIODIC_SOLVE, bcZType = PERIODIC_SOLVE;
    BCKind bcX0 = PERIODIC, bcX1 = PERIODIC, bcY0 = PERIODIC, bcY1 = PERIODIC, bcZ0 = PERIODIC,
           bcZ1 = PERIODIC;
    double periodicDisp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    // [SPONGE]
    SpongeKind spongeKind = RECTILINEAR;
    double spongeP = 1.0 / 1.4, spongeAvgT = 1.0, spongeStrength = 1.0;
    double spongeRectPerc[6] = {0, 0, 0, 0, 0, 0};  // RECT{X0,X1,Y0,Y1,Z0,Z1}PERC
    int spongeCylAxisOrient = 2;
    double spongeCylAxis[3] = {0, 0, 0}, rMin = 0.0;
    // [DOMAIN]
    int Nx = 0, Ny = 0, Nz = 0;
    // [PENCILDECOMP]
    int pRow = 0, pCol = 0;
    // [SOLVER]
    double alphaF = 0.45, mu_ref = 0.0;
    bool useTiming = false;
    FDType xFDType = PADE6, yFDType = PADE6, zFDType = PADE6;
    FilterType filterType = COMPACT10;
    RKType rkType = TVDRK3;
    // [RESTART]
    bool fromRestart = false, onlyGridFromRestart = false, spongeFromRestart = false;
    std::string filename, sponge_filename;
    int timeStep = 0;
    double time = 0.0;
    // [STATS]
    StatsAvgType statsAvgType = NONE;
    bool velocityStats = false, thermoStats = false;
    bool velocityStatsFromRestart = false, thermoStatsFromRestart = false;
    std::string velocityStatsFilename, thermoStatsFilename;
    int stats_interval = 5;
    // [LES]
    StatsAvgType musgsAvgType = NONE;
    LESModel lesModel = NOMODEL;
    bool dumpCoeffRange = false, useTaukk = true;
    int testFilterType = 1;
    // learned closure (no counterpart in the reference): network file written
    // by mc3d_closure_save, and whether out-of-range points fall back to the DSM
    std::string closureFile;
    bool closureFallback = true;
    // the network's inputs: AVERAGED as LESAVERAGING averages the DSM's
    // contractions, or POINT (the features at each point, whatever the averaging
    // of the DSM the network replaces and falls back to)
    bool closurePointFeatures = false;
    double closureFallbackFraction = 1e-3;  // DSM stage when more points than this leave the range

    // Parses INI text.  Returns 0, or a negative code with a message.
    static int parse(const std::string &text, Options &out, std::string &err);
    // The parsed values as "SECTION.KEY = value" lines in the reference's
    // spellings (enums by name, booleans TRUE / FALSE, numbers %.17g).
    std::string describe() const;
};

}  // namespace mc3d
