// INI parsing of the solver options (see options.hpp).
#include "options.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <sstream>

#include "mc3d.h"

namespace mc3d {

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

struct Ini {
    std::map<std::string, std::string> kv;
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    const std::string &get(const std::string &k) const { return kv.at(k); }
};

// boost::program_options config-file syntax: [SECTION] headers, KEY = VALUE,
// '#' comments; keys are reported as SECTION.KEY.
bool parse_ini(const std::string &text, Ini &ini, std::string &err) {
    std::istringstream in(text);
    std::string line, section;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[') {
            if (line.back() != ']') { err = "malformed section header at line " + std::to_string(lineno); return false; }
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) { err = "expected KEY = VALUE at line " + std::to_string(lineno); return false; }
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (key.empty()) { err = "empty key at line " + std::to_string(lineno); return false; }
        const std::string full = section.empty() ? key : section + "." + key;
        if (ini.kv.count(full)) { err = "duplicate key " + full; return false; }
        ini.kv[full] = val;
    }
    return true;
}

struct Reader {
    const Ini &ini;
    std::string &err;
    bool ok = true;

    bool fail(const std::string &m) { if (ok) err = m; ok = false; return false; }

    bool to_double(const std::string &k, double &v) {
        const std::string &s = ini.get(k);
        char *end = nullptr;
        double d = std::strtod(s.c_str(), &end);
        if (s.empty() || *end) return fail("invalid number for " + k + ": '" + s + "'");
        v = d;
        return true;
    }
    bool to_int(const std::string &k, int &v) {
        const std::string &s = ini.get(k);
        char *end = nullptr;
        long d = std::strtol(s.c_str(), &end, 10);
        if (s.empty() || *end) return fail("invalid integer for " + k + ": '" + s + "'");
        v = (int)d;
        return true;
    }
    bool to_bool(const std::string &k, bool &v) {
        std::string s = ini.get(k);
        std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (s == "1" || s == "true" || s == "yes" || s == "on") { v = true; return true; }
        if (s == "0" || s == "false" || s == "no" || s == "off") { v = false; return true; }
        return fail("invalid boolean for " + k + ": '" + ini.get(k) + "'");
    }
    // checkValue: optional with default (src/Options.hpp:602-616)
    template <class T> void opt(const std::string &k, T &v) {
        if (!ini.has(k)) return;
        if constexpr (std::is_same<T, double>::value) to_double(k, v);
        else if constexpr (std::is_same<T, int>::value) to_int(k, v);
        else if constexpr (std::is_same<T, bool>::value) to_bool(k, v);
        else v = ini.get(k);
    }
    // forceValue: mandatory (src/Options.hpp:618-628)
    template <class T> void req(const std::string &k, T &v) {
        if (!ini.has(k)) { fail(k + " is mandatory but not specified"); return; }
        opt(k, v);
    }
    template <class E> void enm(const std::string &k, E &v, std::initializer_list<std::pair<const char *, E>> map) {
        if (!ini.has(k)) { fail(k + " not specified"); return; }
        const std::string &s = ini.get(k);
        for (auto &p : map)
            if (s == p.first) { v = p.second; return; }
        fail("unknown value for " + k + ": '" + s + "'");
    }
};

}  // namespace

int Options::parse(const std::string &text, Options &o, std::string &err) {
    Ini ini;
    if (!parse_ini(text, ini, err)) return MC3D_ECONFIG;
    Reader r{ini, err};
    using O = Options;
    // src/Options.hpp:219-426 in the same order
    r.enm("TIMESTEPPING.TSTYPE", o.TSType, {{"CONST_CFL", O::CONST_CFL}, {"CONST_DT", O::CONST_DT}});
    if (o.TSType == CONST_DT) { r.req("TIMESTEPPING.DT", o.dt); o.CFL = -1.0; }
    else { r.opt("TIMESTEPPING.CFL", o.CFL); o.dt = -1.0; }
    r.opt("TIMESTEPPING.MAXTIME", o.maxTime);
    r.opt("TIMESTEPPING.MAXTIMESTEP", o.maxTimeStep);
    r.opt("TIMESTEPPING.FILTERSTEP", o.filterStep);
    r.opt("TIMESTEPPING.CHECKSTEP", o.checkStep);
    r.opt("TIMESTEPPING.DUMPSTEP", o.dumpStep);
    r.opt("TIMESTEPPING.SUBSTEPFILTERING", o.subStepFiltering);
    r.opt("PENCILDECOMP.PROW", o.pRow);
    r.opt("PENCILDECOMP.PCOL", o.pCol);
    r.opt("SOLVER.ALPHAF", o.alphaF);
    r.req("SOLVER.MU_REF", o.mu_ref);
    r.opt("SOLVER.USETIMING", o.useTiming);
    const std::initializer_list<std::pair<const char *, FDType>> fd = {
        {"CD2", O::CD2}, {"PADE6", O::PADE6}, {"PENTA10", O::PENTA10}};
    r.enm("SOLVER.XFDTYPE", o.xFDType, fd);
    r.enm("SOLVER.YFDTYPE", o.yFDType, fd);
    r.enm("SOLVER.ZFDTYPE", o.zFDType, fd);
    r.enm("SOLVER.FILTERTYPE", o.filterType, {{"COMPACT8", O::COMPACT8}, {"COMPACT10", O::COMPACT10}});
    r.enm("SOLVER.RKTYPE", o.rkType, {{"TVDRK3", O::TVDRK3}, {"RK4", O::RK4}, {"KENRK4", O::KENRK4},
                                       {"LSLDDRK4", O::LSLDDRK4}});
    const std::initializer_list<std::pair<const char *, BCType>> bt = {
        {"PERIODIC_SOLVE", O::PERIODIC_SOLVE}, {"DIRICHLET_SOLVE", O::DIRICHLET_SOLVE}};
    r.enm("BC.BCXTYPE", o.bcXType, bt);
    r.enm("BC.BCYTYPE", o.bcYType, bt);
    r.enm("BC.BCZTYPE", o.bcZType, bt);
    const std::initializer_list<std::pair<const char *, BCKind>> bk = {
        {"INTERNALLY_PERIODIC", O::INTERNALLY_PERIODIC}, {"PERIODIC", O::PERIODIC},
        {"ADIABATIC_WALL", O::ADIABATIC_WALL}, {"SPONGE", O::SPONGE},
        {"CONST_T_WALL", O::CONST_T_WALL}, {"MOVING_ADIABATIC_WALL", O::MOVING_ADIABATIC_WALL},
        {"INLET", O::INLET}};
    r.enm("BC.BCX0", o.bcX0, bk); r.enm("BC.BCX1", o.bcX1, bk);
    r.enm("BC.BCY0", o.bcY0, bk); r.enm("BC.BCY1", o.bcY1, bk);
    r.enm("BC.BCZ0", o.bcZ0, bk); r.enm("BC.BCZ1", o.bcZ1, bk);
    if (!r.ok) return MC3D_ECONFIG;
    // periodic displacements (src/Options.hpp:263-303)
    const char *ax[3] = {"X", "Y", "Z"};
    BCKind lo[3] = {o.bcX0, o.bcY0, o.bcZ0}, hi[3] = {o.bcX1, o.bcY1, o.bcZ1};
    for (int d = 0; d < 3; ++d) {
        if (lo[d] == PERIODIC && hi[d] == PERIODIC) {
            for (int e = 0; e < 3; ++e)
                r.req(std::string("BC.PERIODICDISP") + ax[d] + "_" + ax[e], o.periodicDisp[d][e]);
        } else if ((lo[d] == PERIODIC) != (hi[d] == PERIODIC)) {
            r.fail(std::string("BC mismatch for ") + ax[d] + "0-" + ax[d] + "1 faces");
        }
    }
    // bcValidation (src/Options.cpp:178-268)
    BCType ty[3] = {o.bcXType, o.bcYType, o.bcZType};
    for (int d = 0; d < 3; ++d) {
        auto per = [](BCKind k) { return k == PERIODIC || k == INTERNALLY_PERIODIC; };
        if (ty[d] == PERIODIC_SOLVE) {
            if (!per(lo[d]) || !per(hi[d])) r.fail(std::string("periodic solve in ") + ax[d] + " needs periodic faces");
            if (lo[d] != hi[d]) r.fail(std::string("mixed PERIODIC / INTERNALLY_PERIODIC faces in ") + ax[d]);
        } else if (per(lo[d]) || per(hi[d])) {
            r.fail(std::string("periodic face kinds in a Dirichlet solve of ") + ax[d]);
        }
    }
    // [SPONGE] (src/Options.hpp:313-362): read when a face is a sponge
    {
        const BCKind faces[6] = {o.bcX0, o.bcX1, o.bcY0, o.bcY1, o.bcZ0, o.bcZ1};
        bool any = false;
        for (BCKind f : faces) any = any || f == SPONGE;
        if (any) {
            r.enm("SPONGE.KIND", o.spongeKind, {{"RECTILINEAR", O::RECTILINEAR}, {"CYLINDRICAL", O::CYLINDRICAL}});
            r.opt("SPONGE.AVGTIME", o.spongeAvgT);
            r.opt("SPONGE.PINF", o.spongeP);
            r.opt("SPONGE.STRENGTH", o.spongeStrength);
            if (o.spongeKind == RECTILINEAR) {
                const char *key[6] = {"SPONGE.RECTX0PERC", "SPONGE.RECTX1PERC", "SPONGE.RECTY0PERC",
                                      "SPONGE.RECTY1PERC", "SPONGE.RECTZ0PERC", "SPONGE.RECTZ1PERC"};
                for (int f = 0; f < 6; ++f)
                    if (faces[f] == SPONGE) r.req(key[f], o.spongeRectPerc[f]);
            } else {
                r.req("SPONGE.CYLAXIS_ORIENT", o.spongeCylAxisOrient);
                if (o.spongeCylAxisOrient < 0 || o.spongeCylAxisOrient > 2)
                    r.fail("sponge cylinder orientation must be 0, 1 or 2");
                r.req("SPONGE.CYLAXIS_X", o.spongeCylAxis[0]);
                r.req("SPONGE.CYLAXIS_Y", o.spongeCylAxis[1]);
                r.req("SPONGE.CYLAXIS_Z", o.spongeCylAxis[2]);
                r.req("SPONGE.RMIN", o.rMin);
            }
        }
    }
    r.req("RESTART.FROMRESTART", o.fromRestart);
    r.req("RESTART.ONLYGRIDFROMRESTART", o.onlyGridFromRestart);
    if (o.fromRestart || o.onlyGridFromRestart) r.req("RESTART.FILENAME", o.filename);
    if (o.fromRestart && o.onlyGridFromRestart) r.fail("FROMRESTART and ONLYGRIDFROMRESTART conflict");
    if (!(o.fromRestart || o.onlyGridFromRestart)) {
        r.req("DOMAIN.NX", o.Nx);
        r.req("DOMAIN.NY", o.Ny);
        r.req("DOMAIN.NZ", o.Nz);
    }
    r.req("RESTART.SPONGEAVGFROMRESTART", o.spongeFromRestart);
    if (o.spongeFromRestart) r.req("RESTART.SPONGEAVGFILENAME", o.sponge_filename);
    // src/Options.hpp:394-417
    r.opt("STATS.VELOCITYSTATS", o.velocityStats);
    if (o.velocityStats) {
        r.req("STATS.VELOCITYSTATSFROMRESTART", o.velocityStatsFromRestart);
        if (o.velocityStatsFromRestart) r.req("STATS.VELOCITYSTATSFILENAME", o.velocityStatsFilename);
    }
    r.opt("STATS.THERMOSTATS", o.thermoStats);
    if (o.thermoStats) {
        r.req("STATS.THERMOSTATSFROMRESTART", o.thermoStatsFromRestart);
        if (o.thermoStatsFromRestart) r.req("STATS.THERMOSTATSFILENAME", o.thermoStatsFilename);
    }
    const std::initializer_list<std::pair<const char *, StatsAvgType>> av = {
        {"NONE", O::NONE}, {"XI1_AVG", O::XI1_AVG}, {"XI2_AVG", O::XI2_AVG},
        {"XI3_AVG", O::XI3_AVG}, {"LOCAL", O::LOCAL}, {"GLOBAL", O::GLOBAL}};
    r.enm("STATS.STATSAVGTYPE", o.statsAvgType, av);
    r.opt("STATS.INTERVAL", o.stats_interval);
    if ((o.velocityStats || o.thermoStats) && o.stats_interval <= 0) r.fail("STATS.INTERVAL must be positive");
    // NONE is accepted as an alias of NOMODEL (the reference README spells it so,
    // its parser does not: src/Options.cpp:6-15 vs README.md:47-49)
    r.enm("LES.LESMODEL", o.lesModel, {{"NOMODEL", O::NOMODEL}, {"NONE", O::NOMODEL},
                                       {"VREMAN", O::VREMAN}, {"DSM", O::DSM}, {"LEARNED", O::LEARNED}});
    r.enm("LES.LESAVERAGING", o.musgsAvgType, av);
    r.opt("LES.DSMTESTFILTER", o.testFilterType);
    r.opt("LES.DSMUSETAUKK", o.useTaukk);
    r.opt("LES.DSMCOEFFRANGE", o.dumpCoeffRange);
    r.opt("LES.CLOSUREFILE", o.closureFile);
    r.opt("LES.CLOSUREFALLBACK", o.closureFallback);
    r.opt("LES.CLOSUREFALLBACKFRACTION", o.closureFallbackFraction);
    if (ini.has("LES.CLOSUREFEATURES"))
        r.enm("LES.CLOSUREFEATURES", o.closurePointFeatures, {{"AVERAGED", false}, {"POINT", true}});
    if (!r.ok) return MC3D_ECONFIG;
    // src/Options.hpp:241 (checked after the sizes are known; with a restart
    // they come from the file header and are checked when it is read)
    if (!(o.fromRestart || o.onlyGridFromRestart) && (o.Nx <= 10 || o.Ny <= 10 || o.Nz <= 10)) {
        err = "number of grid points in any direction must be greater than 10";
        return MC3D_ECONFIG;
    }
    if (o.filterStep < 1) { err = "FILTERSTEP must be >= 1"; return MC3D_ECONFIG; }
    if (o.testFilterType != 1 && o.testFilterType != 2) { err = "unknown DSM test filter type"; return MC3D_ECONFIG; }
    return MC3D_OK;
}

std::string Options::describe() const {
    static const char *ts[] = {"CONST_DT", "CONST_CFL"};
    static const char *bt[] = {"PERIODIC_SOLVE", "DIRICHLET_SOLVE"};
    static const char *bk[] = {"PERIODIC", "INTERNALLY_PERIODIC", "SPONGE", "ADIABATIC_WALL", "CONST_T_WALL",
                               "MOVING_ADIABATIC_WALL", "INLET"};
    static const char *sk[] = {"RECTILINEAR", "CYLINDRICAL"};
    static const char *fd[] = {"CD2", "PADE6", "PENTA10"};
    static const char *ft[] = {"COMPACT8", "COMPACT10"};
    static const char *lm[] = {"NOMODEL", "VREMAN", "DSM", "LEARNED"};
    static const char *av[] = {"NONE", "XI1_AVG", "XI2_AVG", "XI3_AVG", "LOCAL", "GLOBAL"};
    static const char *rk[] = {"TVDRK3", "RK4", "KENRK4", "LSLDDRK4"};
    std::string o;
    auto put = [&o](const char *k, const std::string &v) { o += std::string(k) + " = " + v + "\n"; };
    auto num = [](double v) {
        char b[40];
        std::snprintf(b, sizeof b, "%.17g", v);
        return std::string(b);
    };
    auto b = [](bool v) { return std::string(v ? "TRUE" : "FALSE"); };
    auto i = [](int v) { return std::to_string(v); };
    put("TIMESTEPPING.TSTYPE", ts[TSType]);
    put("TIMESTEPPING.CFL", num(CFL));
    put("TIMESTEPPING.DT", num(dt));
    put("TIMESTEPPING.MAXTIME", num(maxTime));
    put("TIMESTEPPING.MAXTIMESTEP", i(maxTimeStep));
    put("TIMESTEPPING.FILTERSTEP", i(filterStep));
    put("TIMESTEPPING.CHECKSTEP", i(checkStep));
    put("TIMESTEPPING.DUMPSTEP", i(dumpStep));
    put("TIMESTEPPING.SUBSTEPFILTERING", b(subStepFiltering));
    put("BC.BCXTYPE", bt[bcXType]);
    put("BC.BCYTYPE", bt[bcYType]);
    put("BC.BCZTYPE", bt[bcZType]);
    const BCKind faces[6] = {bcX0, bcX1, bcY0, bcY1, bcZ0, bcZ1};
    const char *fn[6] = {"BC.BCX0", "BC.BCX1", "BC.BCY0", "BC.BCY1", "BC.BCZ0", "BC.BCZ1"};
    for (int f = 0; f < 6; ++f) put(fn[f], bk[faces[f]]);
    const char *ax = "XYZ";
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
            const std::string k = std::string("BC.PERIODICDISP") + ax[d] + "_" + ax[e];
            put(k.c_str(), num(periodicDisp[d][e]));
        }
    put("SPONGE.KIND", sk[spongeKind]);
    put("SPONGE.STRENGTH", num(spongeStrength));
    put("SPONGE.AVGTIME", num(spongeAvgT));
    put("SPONGE.PINF", num(spongeP));
    put("SPONGE.RMIN", num(rMin));
    put("DOMAIN.NX", i(Nx));
    put("DOMAIN.NY", i(Ny));
    put("DOMAIN.NZ", i(Nz));
    put("PENCILDECOMP.PROW", i(pRow));
    put("PENCILDECOMP.PCOL", i(pCol));
    put("SOLVER.RKTYPE", rk[rkType]);
    put("SOLVER.ALPHAF", num(alphaF));
    put("SOLVER.MU_REF", num(mu_ref));
    put("SOLVER.USETIMING", b(useTiming));
    put("SOLVER.XFDTYPE", fd[xFDType]);
    put("SOLVER.YFDTYPE", fd[yFDType]);
    put("SOLVER.ZFDTYPE", fd[zFDType]);
    put("SOLVER.FILTERTYPE", ft[filterType]);
    put("RESTART.FROMRESTART", b(fromRestart));
    put("RESTART.ONLYGRIDFROMRESTART", b(onlyGridFromRestart));
    put("RESTART.FILENAME", filename);
    put("RESTART.SPONGEAVGFROMRESTART", b(spongeFromRestart));
    put("STATS.INTERVAL", i(stats_interval));
    put("STATS.VELOCITYSTATS", b(velocityStats));
    put("STATS.THERMOSTATS", b(thermoStats));
    put("STATS.STATSAVGTYPE", av[statsAvgType]);
    put("LES.LESMODEL", lm[lesModel]);
    put("LES.LESAVERAGING", av[musgsAvgType]);
    put("LES.DSMTESTFILTER", i(testFilterType));
    put("LES.DSMUSETAUKK", b(useTaukk));
    put("LES.DSMCOEFFRANGE", b(dumpCoeffRange));
    put("LES.CLOSUREFILE", closureFile);
    put("LES.CLOSUREFALLBACK", b(closureFallback));
    put("LES.CLOSUREFEATURES", closurePointFeatures ? "POINT" : "AVERAGED");
    return o;
}

}  // namespace mc3d
