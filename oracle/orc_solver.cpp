// CPU oracle of the CurvilinearCSolver time step (TEST INFRASTRUCTURE ONLY).
// Each block cites the reference lines it restates, including the wall and
// sponge handling of src/CurvilinearCSolver_BC.cpp.
#include "orc_solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

typedef std::vector<double> Vec;

// src/IdealGas.hpp:35-80
struct Gas {
    double gamma = 1.4, Pr = 0.7, p_ref = 1.0 / 1.4, rho_ref = 1.0, T_ref = 1.0, mu_ref = 1.0;
    double R_gas, cp;
    explicit Gas(double mu) : mu_ref(mu) {
        p_ref = 1.0 / gamma;
        R_gas = p_ref / rho_ref / T_ref;
        cp = R_gas * gamma / (gamma - 1.0);
    }
    double solveMu(double T) const { return mu_ref * std::pow(T / T_ref, 0.76); }
    double solveU(double rho, double rhoU) const { return rhoU / rho; }
    double solvep(double rho, double rhoE, double U, double V, double W) const {
        return (gamma - 1) * (rhoE - 0.5 * rho * (U * U + V * V + W * W));
    }
    double solveT(double rho, double p) const { return p / (rho * R_gas); }
    double solveSOS(double rho, double p) const { return std::sqrt(gamma * p / rho); }
    double solverhoE(double rho, double p, double U, double V, double W) const {
        return p / (gamma - 1.0) + 0.5 * rho * (U * U + V * V + W * W);
    }
};

}  // namespace

struct orc_solver {
    orc_config c;
    long n = 0;
    Gas ig{1.0};
    Vec J, J11, J12, J13, J21, J22, J23, J31, J32, J33;
    Vec rho1, rhoU1, rhoV1, rhoW1, rhoE1, rhok, rhoUk, rhoVk, rhoWk, rhoEk;
    Vec rhok2, rhoUk2, rhoVk2, rhoWk2, rhoEk2, rho2, rhoU2, rhoV2, rhoW2, rhoE2;
    Vec U, V, W, T, p, mu, sos, Ucurv, Vcurv, Wcurv;
    Vec dU[3], dV[3], dW[3], dT[3], grad[3][3];
    Vec pre[15], div[15];
    Vec mu_sgs, taukk, k_sgs, C, CI, Prt;
    Vec t1, t2;  // filter temporaries
    // DSM members (src/DSMSGS.hpp:15-27)
    Vec S00, S01, S02, S11, S12, S22, Smag, rho_hat, Smag_hat, rhoU_hat, rhoV_hat, rhoW_hat;
    Vec work[19];
    int timeStep = 0, rkStep = 0, filterTimeStep = 0, dualStep = 1;
    bool rkLast = false;
    double time = 0.0, dt = 0.0, cfl = 0.0;
    // sponge (src/SpongeBC.hpp:38-45)
    bool spongeFlag = false;
    Vec sigma, spgAvg[5];
    // running statistics (src/Stats.hpp:19-40): UAVG VAVG WAVG URMS VRMS WRMS UVREY
    // UWREY VWREY, then RHOAVG RHORMS PAVG PRMS TAVG TRMS
    Vec vstat[9], tstat[6];
    double velWeight = 0.0, thermoWeight = 0.0, statsPrevTime = 0.0;

    // Stats::doAveraging (src/Stats.cpp:248-345): XI1 / XI3 replace phi by its
    // average along the line, summing phi / N in line order; the XI2 branch
    // averages a scratch array and leaves phi as it is (:274-297); the other
    // types do nothing
    void statsAveraging(Vec &phi) {
        const int Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
        if (c.stats_avg == ORC_AVG_XI1) {
            for (int k = 0; k < Nz; ++k)
                for (int j = 0; j < Ny; ++j) {
                    double a = 0.0;
                    for (int i = 0; i < Nx; ++i) a += phi[idx(i, j, k)] / (double)Nx;
                    for (int i = 0; i < Nx; ++i) phi[idx(i, j, k)] = a;
                }
        } else if (c.stats_avg == ORC_AVG_XI3) {
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    double a = 0.0;
                    for (int k = 0; k < Nz; ++k) a += phi[idx(i, j, k)] / (double)Nz;
                    for (int k = 0; k < Nz; ++k) phi[idx(i, j, k)] = a;
                }
        }
    }

    // Stats::updateStatsFields (src/Stats.cpp:121-246): time-weighted running
    // means, RMS and Reynolds stresses of U V W (the current velocities) and of
    // rho2 (the stage register before the filter), p and T
    void updateStats() {
        const double dts = time - statsPrevTime;
        if (c.velocity_stats && dts != 0.0) {
            const double oW = velWeight, tW = oW + dts;
            Vec &UA = vstat[0], &VA = vstat[1], &WA = vstat[2], &UR = vstat[3], &VR = vstat[4], &WR = vstat[5];
            Vec &UV = vstat[6], &UW = vstat[7], &VW = vstat[8];
            for (long ip = 0; ip < n; ++ip) {
                UR[ip] = (oW * (UR[ip] * UR[ip] + UA[ip] * UA[ip]) + dts * U[ip] * U[ip]) / tW;
                VR[ip] = (oW * (VR[ip] * VR[ip] + VA[ip] * VA[ip]) + dts * V[ip] * V[ip]) / tW;
                WR[ip] = (oW * (WR[ip] * WR[ip] + WA[ip] * WA[ip]) + dts * W[ip] * W[ip]) / tW;
                UV[ip] = (oW * (UV[ip] + UA[ip] * VA[ip]) + dts * U[ip] * V[ip]) / tW;
                UW[ip] = (oW * (UW[ip] + UA[ip] * WA[ip]) + dts * U[ip] * W[ip]) / tW;
                VW[ip] = (oW * (VW[ip] + VA[ip] * WA[ip]) + dts * V[ip] * W[ip]) / tW;
                UA[ip] = (oW * UA[ip] + dts * U[ip]) / tW;
                VA[ip] = (oW * VA[ip] + dts * V[ip]) / tW;
                WA[ip] = (oW * WA[ip] + dts * W[ip]) / tW;
            }
            const int order[9] = {3, 4, 5, 6, 7, 8, 0, 1, 2};
            for (int q : order) statsAveraging(vstat[q]);
            for (long ip = 0; ip < n; ++ip) {
                UR[ip] = std::sqrt(std::fabs(UR[ip] - UA[ip] * UA[ip]));
                VR[ip] = std::sqrt(std::fabs(VR[ip] - VA[ip] * VA[ip]));
                WR[ip] = std::sqrt(std::fabs(WR[ip] - WA[ip] * WA[ip]));
                UV[ip] = UV[ip] - UA[ip] * VA[ip];
                UW[ip] = UW[ip] - UA[ip] * WA[ip];
                VW[ip] = VW[ip] - VA[ip] * WA[ip];
            }
            velWeight = tW;
        }
        if (c.thermo_stats && dts != 0.0) {
            const double oW = thermoWeight, tW = oW + dts;
            Vec &RA = tstat[0], &RR = tstat[1], &PA = tstat[2], &PR = tstat[3], &TA = tstat[4], &TR = tstat[5];
            for (long ip = 0; ip < n; ++ip) {
                RR[ip] = (oW * (RR[ip] * RR[ip] + RA[ip] * RA[ip]) + dts * rho2[ip] * rho2[ip]) / tW;
                PR[ip] = (oW * (PR[ip] * PR[ip] + PA[ip] * PA[ip]) + dts * p[ip] * p[ip]) / tW;
                TR[ip] = (oW * (TR[ip] * TR[ip] + TA[ip] * TA[ip]) + dts * T[ip] * T[ip]) / tW;
                RA[ip] = (oW * RA[ip] + dts * rho2[ip]) / tW;
                PA[ip] = (oW * PA[ip] + dts * p[ip]) / tW;
                TA[ip] = (oW * TA[ip] + dts * T[ip]) / tW;
            }
            const int order[6] = {1, 3, 5, 0, 2, 4};
            for (int q : order) statsAveraging(tstat[q]);
            for (long ip = 0; ip < n; ++ip) {
                RR[ip] = std::sqrt(std::fabs(RR[ip] - RA[ip] * RA[ip]));
                PR[ip] = std::sqrt(std::fabs(PR[ip] - PA[ip] * PA[ip]));
                TR[ip] = std::sqrt(std::fabs(TR[ip] - TA[ip] * TA[ip]));
            }
            thermoWeight = tW;
        }
        statsPrevTime = time;
    }

    // ---------------- boundary faces ----------------
    long idx(int i, int j, int k) const { return ((long)k * c.Ny + j) * c.Nx + i; }
    bool is_wall(int f) const {
        return c.face[f] == ORC_FACE_ADIABATIC_WALL || c.face[f] == ORC_FACE_CONST_T_WALL ||
               c.face[f] == ORC_FACE_MOVING_ADIABATIC_WALL;
    }
    // fn(ip, in[6]) for every point of face f (0 X0, 1 X1, 2 Y0, 3 Y1, 4 Z0, 5 Z1);
    // in = the six points inward along the face normal (the FILL_GETMAJIND_YPEN_*
    // index lists of src/Macros.hpp:175-186, of which calcNeumann reads six)
    template <class F>
    void for_face(int f, F fn) {
        const int Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
        const int d = f / 2, hi = f % 2;
        const int n0 = d == 0 ? Ny : Nx, n1 = d == 2 ? Ny : Nz;
        for (int b = 0; b < n1; ++b)
            for (int a = 0; a < n0; ++a) {
                long ip, in[6];
                for (int m = 0; m < 6; ++m) {
                    const int s = hi ? -(m + 1) : (m + 1);
                    if (d == 0) in[m] = idx(hi ? Nx - 1 + s : s, a, b);
                    else if (d == 1) in[m] = idx(a, hi ? Ny - 1 + s : s, b);
                    else in[m] = idx(a, b, hi ? Nz - 1 + s : s);
                }
                if (d == 0) ip = idx(hi ? Nx - 1 : 0, a, b);
                else if (d == 1) ip = idx(a, hi ? Ny - 1 : 0, b);
                else ip = idx(a, b, hi ? Nz - 1 : 0);
                fn(ip, in);
            }
    }
    double neumann_T(int f, const long in[6]) const {
        double v[10] = {};
        for (int m = 0; m < 6; ++m) v[m] = T[in[m]];
        return orc_neumann(c.fd[f / 2], v);
    }
    void curv(long ip) {
        Ucurv[ip] = J11[ip] * U[ip] + J12[ip] * V[ip] + J13[ip] * W[ip];
        Vcurv[ip] = J21[ip] * U[ip] + J22[ip] * V[ip] + J23[ip] * W[ip];
        Wcurv[ip] = J31[ip] * U[ip] + J32[ip] * V[ip] + J33[ip] * W[ip];
    }

    // src/CurvilinearCSolver_BC.cpp:63-322: walls at the initial condition.
    // The static walls come first (X0..Z1), then the moving walls; the wall
    // velocities are zero (src/CurvilinearCSolver.hpp:219-221).  The X1 face
    // applies the Neumann temperature for CONST_T_WALL instead of
    // ADIABATIC_WALL (:103), as the reference does.
    void wallIC() {
        bool any = false;
        for (int f = 0; f < 6; ++f) {
            const int k = c.face[f];
            if (k != ORC_FACE_ADIABATIC_WALL && k != ORC_FACE_CONST_T_WALL) continue;
            any = true;
            const bool neu = f == 1 ? k == ORC_FACE_CONST_T_WALL : k == ORC_FACE_ADIABATIC_WALL;
            for_face(f, [&](long ip, const long in[6]) {
                U[ip] = 0.0; V[ip] = 0.0; W[ip] = 0.0;
                rhoU1[ip] = 0.0; rhoV1[ip] = 0.0; rhoW1[ip] = 0.0;
                if (neu) T[ip] = neumann_T(f, in);
            });
        }
        for (int f = 0; f < 6; ++f) {
            if (c.face[f] != ORC_FACE_MOVING_ADIABATIC_WALL) continue;
            any = true;
            for_face(f, [&](long ip, const long in[6]) {
                U[ip] = 0.0; V[ip] = 0.0; W[ip] = 0.0;
                rhoU1[ip] = rho1[ip] * 0.0; rhoV1[ip] = rho1[ip] * 0.0; rhoW1[ip] = rho1[ip] * 0.0;
                T[ip] = neumann_T(f, in);
            });
        }
        if (!any) return;
        for (long ip = 0; ip < n; ++ip) {  // :311-322
            p[ip] = rho1[ip] * ig.R_gas * T[ip];
            sos[ip] = ig.solveSOS(rho1[ip], p[ip]);
            rhoE1[ip] = ig.solverhoE(rho1[ip], p[ip], U[ip], V[ip], W[ip]);
            curv(ip);
        }
    }

    // src/CurvilinearCSolver_BC.cpp:372-700, at the start of every stage
    void preStepBCHandling() {
        const bool first = rkStep == 1;
        Vec &rhoP = first ? rho1 : rhok, &rhoUP = first ? rhoU1 : rhoUk, &rhoVP = first ? rhoV1 : rhoVk,
            &rhoWP = first ? rhoW1 : rhoWk, &rhoEP = first ? rhoE1 : rhoEk;
        auto fix = [&](long ip) {
            p[ip] = rhoP[ip] * ig.R_gas * T[ip];
            rhoEP[ip] = ig.solverhoE(rhoP[ip], p[ip], U[ip], V[ip], W[ip]);
            curv(ip);
        };
        for (int f = 0; f < 6; ++f) {
            const int k = c.face[f];
            if (k != ORC_FACE_ADIABATIC_WALL && k != ORC_FACE_CONST_T_WALL) continue;
            for_face(f, [&](long ip, const long in[6]) {
                if (f >= 2 && k == ORC_FACE_ADIABATIC_WALL) T[ip] = neumann_T(f, in);
                U[ip] = 0.0; V[ip] = 0.0; W[ip] = 0.0;
                rhoUP[ip] = 0.0; rhoVP[ip] = 0.0; rhoWP[ip] = 0.0;
                if (f < 2 && k == ORC_FACE_ADIABATIC_WALL) T[ip] = neumann_T(f, in);
                fix(ip);
            });
        }
        for (int f = 0; f < 6; ++f) {
            if (c.face[f] != ORC_FACE_MOVING_ADIABATIC_WALL) continue;
            for_face(f, [&](long ip, const long in[6]) {
                if (f >= 2) T[ip] = neumann_T(f, in);
                U[ip] = 0.0; V[ip] = 0.0; W[ip] = 0.0;
                rhoUP[ip] = rhoP[ip] * 0.0; rhoVP[ip] = rhoP[ip] * 0.0; rhoWP[ip] = rhoP[ip] * 0.0;
                if (f < 2) T[ip] = neumann_T(f, in);
                fix(ip);
            });
        }
    }

    // src/CurvilinearCSolver_BC.cpp:702-842, after the k2 assembly: wall faces
    // keep only the normal continuity flux, sponge faces keep nothing
    void postStepBCHandling() {
        for (int f = 0; f < 6; ++f) {
            if (!is_wall(f)) continue;
            const Vec &cont = div[5 * (f / 2)];
            for_face(f, [&](long ip, const long *) {
                rhok2[ip] = -dt * J[ip] * cont[ip];
                rhoUk2[ip] = 0.0; rhoVk2[ip] = 0.0; rhoWk2[ip] = 0.0; rhoEk2[ip] = 0.0;
            });
        }
        for (int f = 0; f < 6; ++f) {
            if (c.face[f] != ORC_FACE_SPONGE) continue;
            for_face(f, [&](long ip, const long *) {
                rhok2[ip] = 0.0; rhoUk2[ip] = 0.0; rhoVk2[ip] = 0.0; rhoWk2[ip] = 0.0; rhoEk2[ip] = 0.0;
            });
        }
    }

    // ---------------- operators ----------------
    void deriv(int dir, const Vec &in, Vec &out) const {
        const double dd = dir == 0 ? c.dx : dir == 1 ? c.dy : c.dz;
        orc_deriv_field(c.fd[dir], dir, c.bc[dir], c.Nx, c.Ny, c.Nz, dd, in.data(), out.data());
    }
    void filt(int kind, double af, int dir, const Vec &in, Vec &out) const {
        orc_filter_field(kind, af, dir, c.bc, c.Nx, c.Ny, c.Nz, in.data(), out.data());
    }
    // filterQuantity / doAveraging(LOCAL) order: Y, then Z, then X.
    // src/DSMSGS.hpp:113-122, src/AbstractSGS.cpp:149-158
    void filterYZX(int kind, double af, const Vec &in, Vec &out) {
        filt(kind, af, 1, in, t1);
        filt(kind, af, 2, t1, t2);
        filt(kind, af, 0, t2, out);
    }
    int sgs_filter_kind() const {
        if (c.les == ORC_VREMAN) return ORC_COMPACT10;
        return c.dsm_test_filter == 2 ? ORC_COMMN5EXP : ORC_COMMN5PADE;
    }
    double sgs_filter_alpha() const { return c.les == ORC_VREMAN ? -0.49 : 0.0; }

    // src/AbstractSGS.cpp:13-164
    void doAveraging(const Vec &phi, Vec &phiF) {
        const int Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
        switch (c.les_avg) {
            case ORC_AVG_LOCAL:
                filterYZX(sgs_filter_kind(), sgs_filter_alpha(), phi, phiF);
                break;
            case ORC_AVG_GLOBAL: {
                double m = 0.0;
                for (long i = 0; i < n; ++i) m += phi[i];
                m /= (double)n;
                for (long i = 0; i < n; ++i) phiF[i] = m;
                break;
            }
            case ORC_AVG_XI1: case ORC_AVG_XI2: case ORC_AVG_XI3: {
                const int dir = c.les_avg - 1;
                const int nl = dir == 0 ? Nx : dir == 1 ? Ny : Nz;
                Vec avg;
                if (dir == 0) {
                    avg.assign((size_t)Ny * Nz, 0.0);
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        avg[(size_t)k * Ny + j] += phi[((size_t)k * Ny + j) * Nx + i] / (double)nl;
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        phiF[((size_t)k * Ny + j) * Nx + i] = avg[(size_t)k * Ny + j];
                } else if (dir == 1) {
                    avg.assign((size_t)Nx * Nz, 0.0);
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        avg[(size_t)k * Nx + i] += phi[((size_t)k * Ny + j) * Nx + i] / (double)nl;
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        phiF[((size_t)k * Ny + j) * Nx + i] = avg[(size_t)k * Nx + i];
                } else {
                    avg.assign((size_t)Nx * Ny, 0.0);
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        avg[(size_t)j * Nx + i] += phi[((size_t)k * Ny + j) * Nx + i] / (double)nl;
                    for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
                        phiF[((size_t)k * Ny + j) * Nx + i] = avg[(size_t)j * Nx + i];
                }
                break;
            }
            default:
                if (&phi != &phiF) phiF = phi;
                break;
        }
    }

    // ---------------- SGS: Vreman, src/VremanSGS.hpp:43-96 ----------------
    void vreman(const Vec &rho) {
        const double csma = 0.17, cv = 2.5 * csma * csma;
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            double vol = c.dx * c.dy * c.dz / J[ip];
            double del = std::pow(vol, 1.0 / 3.0);
            double d1 = del * del, d2 = d1, d3 = d1;
            double d1v1 = grad[0][0][ip], d2v1 = grad[0][1][ip], d3v1 = grad[0][2][ip];
            double d1v2 = grad[1][0][ip], d2v2 = grad[1][1][ip], d3v2 = grad[1][2][ip];
            double d1v3 = grad[2][0][ip], d2v3 = grad[2][1][ip], d3v3 = grad[2][2][ip];
            double b11 = d1 * d1v1 * d1v1 + d2 * d2v1 * d2v1 + d3 * d3v1 * d3v1;
            double b12 = d1 * d1v1 * d1v2 + d2 * d2v1 * d2v2 + d3 * d3v1 * d3v2;
            double b13 = d1 * d1v1 * d1v3 + d2 * d2v1 * d2v3 + d3 * d3v1 * d3v3;
            double b22 = d1 * d1v2 * d1v2 + d2 * d2v2 * d2v2 + d3 * d3v2 * d3v2;
            double b23 = d1 * d1v2 * d1v3 + d2 * d2v2 * d2v3 + d3 * d3v2 * d3v3;
            double b33 = d1 * d1v3 * d1v3 + d2 * d2v3 * d2v3 + d3 * d3v3 * d3v3;
            // d2v3 appears twice and d3v2 not at all (as in the reference)
            double abeta = d1v1 * d1v1 + d1v2 * d1v2 + d1v3 * d1v3 + d2v1 * d2v1 + d2v2 * d2v2 +
                           d2v3 * d2v3 + d3v1 * d3v1 + d2v3 * d2v3 + d3v3 * d3v3;
            double bbeta = b11 * b22 - (b12 * b12) + b11 * b33 - (b13 * b13) + b22 * b33 - (b23 * b23);
            double eps = 1e-12;
            if (bbeta < eps) bbeta = eps;
            double ss = std::sqrt(bbeta / (abeta + eps));
            double m = cv * rho[ip] * ss;
            mu_sgs[ip] = std::min(m, 0.05 * d1 * std::sqrt(2.0 * abeta));
        }
        doAveraging(mu_sgs, mu_sgs);
    }

    // ---------------- SGS: DSM, src/DSMSGS.hpp:124-343 ----------------
    void dsm_mu_taukk(const Vec &rho, const Vec &rhoU, const Vec &rhoV, const Vec &rhoW) {
        const double tf2 = 2.0 * 2.0;  // testFilterRatioSquare, :55
        const int fk = sgs_filter_kind();
        Vec *w = work;  // work[1..18]
        Vec &M00_2 = w[1], &M01_2 = w[2], &M02_2 = w[3], &M11_2 = w[4], &M12_2 = w[5], &M22_2 = w[6];
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            S00[ip] = grad[0][0][ip];
            S11[ip] = grad[1][1][ip];
            S22[ip] = grad[2][2][ip];
            S01[ip] = 0.5 * (grad[0][1][ip] + grad[1][0][ip]);
            S02[ip] = 0.5 * (grad[0][2][ip] + grad[2][0][ip]);
            S12[ip] = 0.5 * (grad[1][2][ip] + grad[2][1][ip]);
            Smag[ip] = std::sqrt(2.0 * (S00[ip] * S00[ip] + S11[ip] * S11[ip] + S22[ip] * S22[ip]) +
                                 4.0 * (S01[ip] * S01[ip] + S02[ip] * S02[ip] + S12[ip] * S12[ip]));
            M00_2[ip] = 2.0 * rho[ip] * Smag[ip] * (S00[ip] - (1.0 / 3.0) * (S00[ip] + S11[ip] + S22[ip]));
            M11_2[ip] = 2.0 * rho[ip] * Smag[ip] * (S11[ip] - (1.0 / 3.0) * (S00[ip] + S11[ip] + S22[ip]));
            M22_2[ip] = 2.0 * rho[ip] * Smag[ip] * (S22[ip] - (1.0 / 3.0) * (S00[ip] + S11[ip] + S22[ip]));
            M01_2[ip] = 2.0 * rho[ip] * Smag[ip] * S01[ip];
            M02_2[ip] = 2.0 * rho[ip] * Smag[ip] * S02[ip];
            M12_2[ip] = 2.0 * rho[ip] * Smag[ip] * S12[ip];
        }
        Vec &M00 = w[7], &M01 = w[8], &M02 = w[9], &M11 = w[10], &M12 = w[11], &M22 = w[12];
        filterYZX(fk, 0.0, M00_2, M00);
        filterYZX(fk, 0.0, M01_2, M01);
        filterYZX(fk, 0.0, M02_2, M02);
        filterYZX(fk, 0.0, M11_2, M11);
        filterYZX(fk, 0.0, M12_2, M12);
        filterYZX(fk, 0.0, M22_2, M22);
        Vec &S00_hat = w[1], &S01_hat = w[2], &S02_hat = w[3], &S11_hat = w[4], &S12_hat = w[5],
            &S22_hat = w[6];
        filterYZX(fk, 0.0, rho, rho_hat);
        filterYZX(fk, 0.0, S00, S00_hat);
        filterYZX(fk, 0.0, S01, S01_hat);
        filterYZX(fk, 0.0, S02, S02_hat);
        filterYZX(fk, 0.0, S11, S11_hat);
        filterYZX(fk, 0.0, S12, S12_hat);
        filterYZX(fk, 0.0, S22, S22_hat);
        filterYZX(fk, 0.0, Smag, Smag_hat);
        Vec &L00_t = w[13], &L01_t = w[14], &L02_t = w[15], &L11_t = w[16], &L12_t = w[17], &L22_t = w[18];
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            M00[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * (S00_hat[ip] - (1.0 / 3.0) * (S00_hat[ip] + S11_hat[ip] + S22_hat[ip]));
            M11[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * (S11_hat[ip] - (1.0 / 3.0) * (S00_hat[ip] + S11_hat[ip] + S22_hat[ip]));
            M22[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * (S22_hat[ip] - (1.0 / 3.0) * (S00_hat[ip] + S11_hat[ip] + S22_hat[ip]));
            M01[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * S01_hat[ip];
            M02[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * S02_hat[ip];
            M12[ip] -= 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * S12_hat[ip];
            L00_t[ip] = rhoU[ip] * rhoU[ip] / rho[ip];
            L01_t[ip] = rhoU[ip] * rhoV[ip] / rho[ip];
            L02_t[ip] = rhoU[ip] * rhoW[ip] / rho[ip];
            L11_t[ip] = rhoV[ip] * rhoV[ip] / rho[ip];
            L12_t[ip] = rhoV[ip] * rhoW[ip] / rho[ip];
            L22_t[ip] = rhoW[ip] * rhoW[ip] / rho[ip];
        }
        Vec &L00 = w[1], &L01 = w[2], &L02 = w[3], &L11 = w[4], &L12 = w[5], &L22 = w[6];
        filterYZX(fk, 0.0, L00_t, L00);
        filterYZX(fk, 0.0, L01_t, L01);
        filterYZX(fk, 0.0, L02_t, L02);
        filterYZX(fk, 0.0, L11_t, L11);
        filterYZX(fk, 0.0, L12_t, L12);
        filterYZX(fk, 0.0, L22_t, L22);
        filterYZX(fk, 0.0, rhoU, rhoU_hat);
        filterYZX(fk, 0.0, rhoV, rhoV_hat);
        filterYZX(fk, 0.0, rhoW, rhoW_hat);
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            L00[ip] -= rhoU_hat[ip] * rhoU_hat[ip] / rho_hat[ip];
            L01[ip] -= rhoU_hat[ip] * rhoV_hat[ip] / rho_hat[ip];
            L02[ip] -= rhoU_hat[ip] * rhoW_hat[ip] / rho_hat[ip];
            L11[ip] -= rhoV_hat[ip] * rhoV_hat[ip] / rho_hat[ip];
            L12[ip] -= rhoV_hat[ip] * rhoW_hat[ip] / rho_hat[ip];
            L22[ip] -= rhoW_hat[ip] * rhoW_hat[ip] / rho_hat[ip];
        }
        Vec &CIdenom = w[13], &alpha_hat = w[14];
        if (c.dsm_use_taukk) {
#pragma omp parallel for
            for (long ip = 0; ip < n; ++ip) CIdenom[ip] = 2.0 * rho[ip] * Smag[ip] * Smag[ip];
            filterYZX(fk, 0.0, CIdenom, alpha_hat);
#pragma omp parallel for
            for (long ip = 0; ip < n; ++ip)
                CIdenom[ip] = 2.0 * tf2 * rho_hat[ip] * Smag_hat[ip] * Smag_hat[ip] - alpha_hat[ip];
        }
        Vec &Lkk = w[15], &LijMij = w[17], &MijMij = w[18];
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            Lkk[ip] = L00[ip] + L11[ip] + L22[ip];
            LijMij[ip] = (L00[ip] - (1.0 / 3.0) * Lkk[ip]) * M00[ip] + (L11[ip] - (1.0 / 3.0) * Lkk[ip]) * M11[ip] +
                         (L22[ip] - (1.0 / 3.0) * Lkk[ip]) * M22[ip] + 2.0 * L01[ip] * M01[ip] +
                         2.0 * L02[ip] * M02[ip] + 2.0 * L12[ip] * M12[ip];
            MijMij[ip] = M00[ip] * M00[ip] + M11[ip] * M11[ip] + M22[ip] * M22[ip] + 2.0 * M01[ip] * M01[ip] +
                         2.0 * M02[ip] * M02[ip] + 2.0 * M12[ip] * M12[ip];
        }
        Vec &Lkk_avg = w[7], &CIdenom_avg = w[8], &LijMij_avg = w[9], &MijMij_avg = w[10];
        if (c.dsm_use_taukk) {
            doAveraging(Lkk, Lkk_avg);
            doAveraging(CIdenom, CIdenom_avg);
        }
        doAveraging(LijMij, LijMij_avg);
        doAveraging(MijMij, MijMij_avg);
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            mu_sgs[ip] = rho[ip] * Smag[ip] * LijMij_avg[ip] / MijMij_avg[ip];
            taukk[ip] = c.dsm_use_taukk ? 2.0 * rho[ip] * Smag[ip] * Smag[ip] * Lkk_avg[ip] / CIdenom_avg[ip] : 0.0;
        }
        if (c.dsm_coeff_range) {
#pragma omp parallel for
            for (long ip = 0; ip < n; ++ip) {
                double vol = c.dx * c.dy * c.dz / J[ip];
                double del = std::pow(vol, 1.0 / 3.0);
                C[ip] = LijMij_avg[ip] / (del * del * MijMij_avg[ip]);
                CI[ip] = c.dsm_use_taukk ? Lkk_avg[ip] / (del * del * CIdenom_avg[ip]) : 0.0;
            }
        }
    }

    // src/DSMSGS.hpp:345-444
    void dsm_ksgs(const Vec gT[3], const Vec &rho, const Vec &rhoU, const Vec &rhoV, const Vec &rhoW,
                  const Vec &Tm) {
        const double tf2 = 2.0 * 2.0;
        const int fk = sgs_filter_kind();
        Vec *w = work;
        Vec &N_0 = w[1], &N_1 = w[2], &N_2 = w[3], &NRHS_0 = w[4], &NRHS_1 = w[5], &NRHS_2 = w[6];
        Vec &K_0 = w[7], &K_1 = w[8], &K_2 = w[9], &KRHS_0 = w[10], &KRHS_1 = w[11], &KRHS_2 = w[12];
        Vec &g0h = w[13], &g1h = w[14], &g2h = w[15], &rhoT = w[16], &rhoT_hat = w[17];
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            N_0[ip] = rho[ip] * Smag[ip] * gT[0][ip];
            N_1[ip] = rho[ip] * Smag[ip] * gT[1][ip];
            N_2[ip] = rho[ip] * Smag[ip] * gT[2][ip];
            K_0[ip] = rhoU[ip] * Tm[ip];
            K_1[ip] = rhoV[ip] * Tm[ip];
            K_2[ip] = rhoW[ip] * Tm[ip];
            rhoT[ip] = rho[ip] * Tm[ip];
        }
        filterYZX(fk, 0.0, N_0, NRHS_0);
        filterYZX(fk, 0.0, N_1, NRHS_1);
        filterYZX(fk, 0.0, N_2, NRHS_2);
        filterYZX(fk, 0.0, K_0, KRHS_0);
        filterYZX(fk, 0.0, K_1, KRHS_1);
        filterYZX(fk, 0.0, K_2, KRHS_2);
        filterYZX(fk, 0.0, rhoT, rhoT_hat);
        filterYZX(fk, 0.0, gT[0], g0h);
        filterYZX(fk, 0.0, gT[1], g1h);
        filterYZX(fk, 0.0, gT[2], g2h);
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            N_0[ip] = tf2 * rho_hat[ip] * Smag_hat[ip] * g0h[ip] - NRHS_0[ip];
            N_1[ip] = tf2 * rho_hat[ip] * Smag_hat[ip] * g1h[ip] - NRHS_1[ip];
            N_2[ip] = tf2 * rho_hat[ip] * Smag_hat[ip] * g2h[ip] - NRHS_2[ip];
            // K_i divides by rho, not rho_hat (as in the reference)
            K_0[ip] = -rhoU_hat[ip] * rhoT_hat[ip] / rho[ip] + KRHS_0[ip];
            K_1[ip] = -rhoV_hat[ip] * rhoT_hat[ip] / rho[ip] + KRHS_1[ip];
            K_2[ip] = -rhoW_hat[ip] * rhoT_hat[ip] / rho[ip] + KRHS_2[ip];
        }
        Vec &NiNi = w[4], &NiKi = w[5];
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            NiNi[ip] = N_0[ip] * N_0[ip] + N_1[ip] * N_1[ip] + N_2[ip] * N_2[ip];
            NiKi[ip] = K_0[ip] * N_0[ip] + K_1[ip] * N_1[ip] + K_2[ip] * N_2[ip];
        }
        Vec &NiNi_avg = w[6], &NiKi_avg = w[10];
        doAveraging(NiNi, NiNi_avg);
        doAveraging(NiKi, NiKi_avg);
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            double C_over_Prt = NiKi_avg[ip] / NiNi_avg[ip];
            k_sgs[ip] = -ig.cp * C_over_Prt * rho[ip] * Smag[ip];
        }
        if (c.dsm_coeff_range) {
#pragma omp parallel for
            for (long ip = 0; ip < n; ++ip) {
                double vol = c.dx * c.dy * c.dz / J[ip];
                double del = std::pow(vol, 1.0 / 3.0);
                Prt[ip] = -C[ip] * del * del * NiNi_avg[ip] / NiKi_avg[ip];
            }
        }
    }

    // ---------------- solver core ----------------
    // src/CurvilinearCSolver_Core.cpp:189-237
    void calcDtFromCFL() {
        double lmax = -100000.0;
        for (long ip = 0; ip < n; ++ip) {
            double v = J11[ip] * (std::fabs(U[ip]) + sos[ip]) / c.dx +
                       J22[ip] * (std::fabs(V[ip]) + sos[ip]) / c.dy +
                       J33[ip] * (std::fabs(W[ip]) + sos[ip]) / c.dz;
            if (v > lmax) lmax = v;
        }
        if (c.ts_type == ORC_CONST_CFL) dt = c.cfl / lmax;
        else cfl = dt * lmax;
        if (timeStep == c.timestep0) { timeStep++; time = c.time0; }
        else { timeStep++; time += dt; }
    }

    // gradU[i][j] = d u_i / d x_j, src/CurvilinearCSolver_Core.cpp:548-561
    void gradient_tensor(Vec d[3], Vec g[3]) {
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            g[0][ip] = J11[ip] * d[0][ip] + J21[ip] * d[1][ip] + J31[ip] * d[2][ip];
            g[1][ip] = J12[ip] * d[0][ip] + J22[ip] * d[1][ip] + J32[ip] * d[2][ip];
            g[2][ip] = J13[ip] * d[0][ip] + J23[ip] * d[1][ip] + J33[ip] * d[2][ip];
        }
    }

    // src/CurvilinearCSolver_Core.cpp:468-787
    void preStepDerivatives() {
        const bool first = rkStep == 1;
        const Vec &rhoP = first ? rho1 : rhok, &rhoUP = first ? rhoU1 : rhoUk,
                  &rhoVP = first ? rhoV1 : rhoVk, &rhoWP = first ? rhoW1 : rhoWk,
                  &rhoEP = first ? rhoE1 : rhoEk;
        for (int d = 0; d < 3; ++d) { deriv(d, U, dU[d]); deriv(d, V, dV[d]); deriv(d, W, dW[d]); }
        gradient_tensor(dU, grad[0]);
        gradient_tensor(dV, grad[1]);
        gradient_tensor(dW, grad[2]);
        if (c.les == ORC_VREMAN) {
            vreman(rhoP);
        } else if (c.les == ORC_DSM) {
            dsm_mu_taukk(rhoP, rhoUP, rhoVP, rhoWP);
            // taukk folded into pressure before grad T, :573-585
#pragma omp parallel for
            for (long ip = 0; ip < n; ++ip) {
                p[ip] -= 0.5 * taukk[ip];
                T[ip] = ig.solveT(rhoP[ip], p[ip]);
                mu[ip] = ig.solveMu(T[ip]);
                sos[ip] = ig.solveSOS(rhoP[ip], p[ip]);
            }
        }
        for (int d = 0; d < 3; ++d) deriv(d, T, dT[d]);
        if (c.les == ORC_DSM) {
            Vec gT[3] = {Vec(n), Vec(n), Vec(n)};
            gradient_tensor(dT, gT);
            dsm_ksgs(gT, rhoP, rhoUP, rhoVP, rhoWP, T);
        }
        const double Pr_t = 0.9;
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            double dUdx = grad[0][0][ip], dUdy = grad[0][1][ip], dUdz = grad[0][2][ip];
            double dVdx = grad[1][0][ip], dVdy = grad[1][1][ip], dVdz = grad[1][2][ip];
            double dWdx = grad[2][0][ip], dWdy = grad[2][1][ip], dWdz = grad[2][2][ip];
            double T11 = (4.0 / 3.0) * dUdx - (2.0 / 3.0) * dVdy - (2.0 / 3.0) * dWdz;
            double T22 = -(2.0 / 3.0) * dUdx + (4.0 / 3.0) * dVdy - (2.0 / 3.0) * dWdz;
            double T33 = -(2.0 / 3.0) * dUdx - (2.0 / 3.0) * dVdy + (4.0 / 3.0) * dWdz;
            double T12 = dUdy + dVdx, T13 = dUdz + dWdx, T23 = dVdz + dWdy;
            double mu_eff, k_eff;
            if (c.les != ORC_NOMODEL) {
                mu_eff = mu[ip] + mu_sgs[ip];
                if (c.les == ORC_DSM) k_eff = ig.cp * mu[ip] / ig.Pr + k_sgs[ip];
                else k_eff = ig.cp * (mu[ip] / ig.Pr + mu_sgs[ip] / Pr_t);
                if (mu_eff < 0.0) mu_eff = 0.0;
                if (k_eff < 0.0) k_eff = 0.0;
            } else {
                mu_eff = mu[ip];
                k_eff = (ig.cp / ig.Pr) * mu[ip];
            }
            T11 *= mu_eff; T22 *= mu_eff; T33 *= mu_eff;
            T12 *= mu_eff; T13 *= mu_eff; T23 *= mu_eff;
            const double T21 = T12, T31 = T13, T32 = T23;
            double q1 = k_eff * (J11[ip] * dT[0][ip] + J21[ip] * dT[1][ip] + J31[ip] * dT[2][ip]);
            double q2 = k_eff * (J12[ip] * dT[0][ip] + J22[ip] * dT[1][ip] + J32[ip] * dT[2][ip]);
            double q3 = k_eff * (J13[ip] * dT[0][ip] + J23[ip] * dT[1][ip] + J33[ip] * dT[2][ip]);
            double b1 = U[ip] * T11 + V[ip] * T12 + W[ip] * T13 + q1;
            double b2 = U[ip] * T21 + V[ip] * T22 + W[ip] * T23 + q2;
            double b3 = U[ip] * T31 + V[ip] * T32 + W[ip] * T33 + q3;
            const double iJ = 1.0 / J[ip];
            pre[0][ip] = iJ * (rhoP[ip] * Ucurv[ip]);
            pre[5][ip] = iJ * (rhoP[ip] * Vcurv[ip]);
            pre[10][ip] = iJ * (rhoP[ip] * Wcurv[ip]);
            double F1, F2, F3, G1, G2, G3;
            F1 = rhoUP[ip] * Ucurv[ip] + J11[ip] * p[ip];
            F2 = rhoUP[ip] * Vcurv[ip] + J21[ip] * p[ip];
            F3 = rhoUP[ip] * Wcurv[ip] + J31[ip] * p[ip];
            G1 = J11[ip] * T11 + J12[ip] * T21 + J13[ip] * T31;
            G2 = J21[ip] * T11 + J22[ip] * T21 + J23[ip] * T31;
            G3 = J31[ip] * T11 + J32[ip] * T21 + J33[ip] * T31;
            pre[1][ip] = iJ * (F1 - G1); pre[6][ip] = iJ * (F2 - G2); pre[11][ip] = iJ * (F3 - G3);
            F1 = rhoVP[ip] * Ucurv[ip] + J12[ip] * p[ip];
            F2 = rhoVP[ip] * Vcurv[ip] + J22[ip] * p[ip];
            F3 = rhoVP[ip] * Wcurv[ip] + J32[ip] * p[ip];
            G1 = J11[ip] * T12 + J12[ip] * T22 + J13[ip] * T32;
            G2 = J21[ip] * T12 + J22[ip] * T22 + J23[ip] * T32;
            G3 = J31[ip] * T12 + J32[ip] * T22 + J33[ip] * T32;
            pre[2][ip] = iJ * (F1 - G1); pre[7][ip] = iJ * (F2 - G2); pre[12][ip] = iJ * (F3 - G3);
            F1 = rhoWP[ip] * Ucurv[ip] + J13[ip] * p[ip];
            F2 = rhoWP[ip] * Vcurv[ip] + J23[ip] * p[ip];
            F3 = rhoWP[ip] * Wcurv[ip] + J33[ip] * p[ip];
            G1 = J11[ip] * T13 + J12[ip] * T23 + J13[ip] * T33;
            G2 = J21[ip] * T13 + J22[ip] * T23 + J23[ip] * T33;
            G3 = J31[ip] * T13 + J32[ip] * T23 + J33[ip] * T33;
            pre[3][ip] = iJ * (F1 - G1); pre[8][ip] = iJ * (F2 - G2); pre[13][ip] = iJ * (F3 - G3);
            F1 = (rhoEP[ip] + p[ip]) * Ucurv[ip];
            F2 = (rhoEP[ip] + p[ip]) * Vcurv[ip];
            F3 = (rhoEP[ip] + p[ip]) * Wcurv[ip];
            G1 = J11[ip] * b1 + J12[ip] * b2 + J13[ip] * b3;
            G2 = J21[ip] * b1 + J22[ip] * b2 + J23[ip] * b3;
            G3 = J31[ip] * b1 + J32[ip] * b2 + J33[ip] * b3;
            pre[4][ip] = iJ * (F1 - G1); pre[9][ip] = iJ * (F2 - G2); pre[14][ip] = iJ * (F3 - G3);
        }
        // src/CurvilinearCSolver_Core.cpp:763-773: xi of the _1 set, eta of _2, zeta of _3
        for (int e = 0; e < 5; ++e) {
            deriv(0, pre[e], div[e]);
            deriv(1, pre[5 + e], div[5 + e]);
            deriv(2, pre[10 + e], div[10 + e]);
        }
    }

    // src/CurvilinearCSolver_Core.cpp:790-947 (sources are zero; the momentum
    // increments accumulate into the previous k2, as the reference writes them)
    void solveEqnSet() {
        const bool first = rkStep == 1;
        const Vec *q[5] = {first ? &rho1 : &rhok, first ? &rhoU1 : &rhoUk, first ? &rhoV1 : &rhoVk,
                           first ? &rhoW1 : &rhoWk, first ? &rhoE1 : &rhoEk};
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            // calcSpongeSource = sigma (avg - phi), src/CurvilinearCSolver.hpp:293-295
            double z[5];
            for (int e = 0; e < 5; ++e) {
                const double spg = spongeFlag ? sigma[ip] * (spgAvg[e][ip] - (*q[e])[ip]) : 0.0;
                z[e] = (spg + 0.0) / J[ip];
            }
            rhok2[ip] = dt * J[ip] * (-div[0][ip] - div[5][ip] - div[10][ip] + z[0]);
            rhoUk2[ip] += -div[1][ip] - div[6][ip] - div[11][ip] + z[1];
            rhoUk2[ip] *= dt * J[ip];
            rhoVk2[ip] += -div[2][ip] - div[7][ip] - div[12][ip] + z[2];
            rhoVk2[ip] *= dt * J[ip];
            rhoWk2[ip] += -div[3][ip] - div[8][ip] - div[13][ip] + z[3];
            rhoWk2[ip] *= dt * J[ip];
            rhoEk2[ip] = -div[4][ip] - div[9][ip] - div[14][ip] + z[4];
            rhoEk2[ip] *= dt * J[ip];
        }
    }

    // RK register updates: TVDRK3 src/TVDRK3.hpp:50-133, RK4 src/RK4.hpp:47-125,
    // KenRK4 src/KenRK4.hpp:53-152, LSLDDRK4 src/LSLDDRK4.hpp:93-322.
    void updateConservedData() {
        Vec *r1[5] = {&rho1, &rhoU1, &rhoV1, &rhoW1, &rhoE1};
        Vec *rk[5] = {&rhok, &rhoUk, &rhoVk, &rhoWk, &rhoEk};
        Vec *r2k[5] = {&rhok2, &rhoUk2, &rhoVk2, &rhoWk2, &rhoEk2};
        Vec *r2[5] = {&rho2, &rhoU2, &rhoV2, &rhoW2, &rhoE2};
        const int s = rkStep;
        for (int q = 0; q < 5; ++q) {
            Vec &a1 = *r1[q], &ak = *rk[q], &k2 = *r2k[q], &a2 = *r2[q];
            if (c.rk == ORC_TVDRK3) {
                if (s == 1) for (long i = 0; i < n; ++i) ak[i] = a1[i] + 1.0 * k2[i];
                else if (s == 2) for (long i = 0; i < n; ++i) ak[i] = (3.0 / 4.0) * a1[i] + (1.0 / 4.0) * ak[i] + (1.0 / 4.0) * k2[i];
                else for (long i = 0; i < n; ++i) a2[i] = (1.0 / 3.0) * a1[i] + (2.0 / 3.0) * ak[i] + (2.0 / 3.0) * k2[i];
            } else if (c.rk == ORC_RK4) {
                const double a[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
                const double b[3] = {1.0 / 2.0, 1.0 / 2.0, 1.0};
                if (s == 1) for (long i = 0; i < n; ++i) { a2[i] = a1[i] + a[0] * k2[i]; ak[i] = a1[i] + b[0] * k2[i]; }
                else if (s < 4) for (long i = 0; i < n; ++i) { a2[i] += a[s - 1] * k2[i]; ak[i] = a1[i] + b[s - 1] * k2[i]; }
                else for (long i = 0; i < n; ++i) a2[i] += a[3] * k2[i];
            } else {
                double A[6], B[6];
                int S;
                lowstorage_coeffs(A, B, S);
                if (s == 1) {
                    for (long i = 0; i < n; ++i) { ak[i] = a1[i] + B[0] * k2[i]; a1[i] = k2[i]; }
                } else if (s < S) {
                    for (long i = 0; i < n; ++i) { a1[i] = k2[i] + A[s - 1] * a1[i]; ak[i] += B[s - 1] * a1[i]; }
                } else {
                    for (long i = 0; i < n; ++i) { a1[i] = k2[i] + A[s - 1] * a1[i]; a2[i] = ak[i] + B[s - 1] * a1[i]; }
                }
            }
        }
    }

    void lowstorage_coeffs(double *A, double *B, int &S) const {
        if (c.rk == ORC_KENRK4) {  // src/KenRK4.hpp:30-40
            const double a[5] = {0.0, -6234157559845.0 / 12983515589748.0, -6194124222391.0 / 4410992767914.0,
                                 -31623096876824.0 / 15682348800105.0, -12251185447671.0 / 11596622555746.0};
            const double b[5] = {494393426753.0 / 4806282396855.0, 4047970641027.0 / 5463924506627.0,
                                 9795748752853.0 / 13190207949281.0, 4009051133189.0 / 8539092990294.0,
                                 1348533437543.0 / 7166442652324.0};
            for (int i = 0; i < 5; ++i) { A[i] = a[i]; B[i] = b[i]; }
            S = 5;
        } else if (dualStep == 1) {  // src/LSLDDRK4.hpp:40-56
            const double a[5] = {0.0, -0.60512264332862261228, -2.04375640234761394333,
                                 -0.74069990637544192841, -4.42317651302968168941};
            const double b[5] = {0.26874543888713438496, 0.8014706973220802933, 0.50515704269422722538,
                                 0.56235680379000296407, 0.05900655127758823335};
            for (int i = 0; i < 5; ++i) { A[i] = a[i]; B[i] = b[i]; }
            S = 5;
        } else {  // src/LSLDDRK4.hpp:59-78
            const double a[6] = {0.0, -0.44127377153877382565, -1.073982008079781868, -1.7063570791256758809,
                                 -2.7979293162682443056, -4.0913537120919160454};
            const double b[6] = {0.11584888181285561688, 0.37287699051652864918, 0.73795368921435295698,
                                 0.57981109366311039538, 1.031284991300145194, 0.15};
            for (int i = 0; i < 6; ++i) { A[i] = a[i]; B[i] = b[i]; }
            S = 6;
        }
    }

    int n_stages() const {
        if (c.rk == ORC_TVDRK3) return 3;
        if (c.rk == ORC_RK4) return 4;
        if (c.rk == ORC_KENRK4) return 5;
        return dualStep == 1 ? 5 : 6;
    }

    // src/CurvilinearCSolver_Core.cpp:949-1155 (round-robin direction order)
    void filterConservedData() {
        Vec *in[5], *out[5];
        if (!rkLast) {
            Vec *k[5] = {&rhok, &rhoUk, &rhoVk, &rhoWk, &rhoEk};
            for (int q = 0; q < 5; ++q) in[q] = out[q] = k[q];
        } else {
            Vec *a2[5] = {&rho2, &rhoU2, &rhoV2, &rhoW2, &rhoE2};
            Vec *a1[5] = {&rho1, &rhoU1, &rhoV1, &rhoW1, &rhoE1};
            for (int q = 0; q < 5; ++q) { in[q] = a2[q]; out[q] = a1[q]; }
        }
        if (timeStep % c.filter_step == 0) {
            filterTimeStep++;
            int order[3];
            if (filterTimeStep % 3 == 1) { order[0] = 1; order[1] = 2; order[2] = 0; }
            else if (filterTimeStep % 3 == 2) { order[0] = 2; order[1] = 0; order[2] = 1; }
            else { order[0] = 0; order[1] = 1; order[2] = 2; }
            for (int q = 0; q < 5; ++q) {
                Vec a(n), b(n);
                filt(c.filter, c.alphaF, order[0], *in[q], a);
                filt(c.filter, c.alphaF, order[1], a, b);
                filt(c.filter, c.alphaF, order[2], b, *out[q]);
            }
        } else if (rkLast) {
            rho1 = rho2; rhoU1 = rhoU2; rhoV1 = rhoV2; rhoW1 = rhoW2; rhoE1 = rhoE2;
        }
    }

    // src/CurvilinearCSolver_Core.cpp:1158-1198
    void updateNonConservedData() {
        const Vec &r = rkLast ? rho1 : rhok, &ru = rkLast ? rhoU1 : rhoUk, &rv = rkLast ? rhoV1 : rhoVk,
                  &rw = rkLast ? rhoW1 : rhoWk, &re = rkLast ? rhoE1 : rhoEk;
#pragma omp parallel for
        for (long ip = 0; ip < n; ++ip) {
            U[ip] = ig.solveU(r[ip], ru[ip]);
            V[ip] = ig.solveU(r[ip], rv[ip]);
            W[ip] = ig.solveU(r[ip], rw[ip]);
            p[ip] = ig.solvep(r[ip], re[ip], U[ip], V[ip], W[ip]);
            T[ip] = ig.solveT(r[ip], p[ip]);
            mu[ip] = ig.solveMu(T[ip]);
            sos[ip] = ig.solveSOS(r[ip], p[ip]);
            Ucurv[ip] = J11[ip] * U[ip] + J12[ip] * V[ip] + J13[ip] * W[ip];
            Vcurv[ip] = J21[ip] * U[ip] + J22[ip] * V[ip] + J23[ip] * W[ip];
            Wcurv[ip] = J31[ip] * U[ip] + J32[ip] * V[ip] + J33[ip] * W[ip];
        }
    }

    // one executeSolverLoop iteration, src/TVDRK3.hpp:135-176 and its siblings
    void step() {
        rkLast = false;
        calcDtFromCFL();
        const int S = n_stages();
        for (int s = 1; s <= S; ++s) {
            rkStep = s;
            if (s == S) rkLast = true;
            preStepBCHandling();
            preStepDerivatives();
            solveEqnSet();
            postStepBCHandling();
            updateConservedData();
            if (rkLast || c.substep_filtering) filterConservedData();
            updateNonConservedData();
        }
        if (c.rk == ORC_LSLDDRK4) dualStep = dualStep == 1 ? 2 : 1;
        // postStep, src/CurvilinearCSolver.hpp:355-365
        if ((c.velocity_stats || c.thermo_stats) && c.stats_interval > 0 && timeStep % c.stats_interval == 0)
            updateStats();
    }
};

extern "C" {

orc_solver *orc_solver_create(const orc_config *cfg, const double *const metrics[10],
                              const double *rho0, const double *U0, const double *V0,
                              const double *W0, const double *p0) {
    auto *s = new orc_solver();
    s->c = *cfg;
    s->n = (long)cfg->Nx * cfg->Ny * cfg->Nz;
    s->ig = Gas(cfg->mu_ref);
    const long n = s->n;
    Vec *m[10] = {&s->J, &s->J11, &s->J12, &s->J13, &s->J21, &s->J22, &s->J23, &s->J31, &s->J32, &s->J33};
    for (int k = 0; k < 10; ++k) m[k]->assign(metrics[k], metrics[k] + n);
    Vec *all[] = {&s->rho1, &s->rhoU1, &s->rhoV1, &s->rhoW1, &s->rhoE1, &s->rhok, &s->rhoUk, &s->rhoVk,
                  &s->rhoWk, &s->rhoEk, &s->rhok2, &s->rhoUk2, &s->rhoVk2, &s->rhoWk2, &s->rhoEk2,
                  &s->rho2, &s->rhoU2, &s->rhoV2, &s->rhoW2, &s->rhoE2, &s->U, &s->V, &s->W, &s->T,
                  &s->p, &s->mu, &s->sos, &s->Ucurv, &s->Vcurv, &s->Wcurv, &s->mu_sgs, &s->taukk,
                  &s->k_sgs, &s->C, &s->CI, &s->Prt, &s->t1, &s->t2};
    for (Vec *v : all) v->assign(n, 0.0);  // k2 registers start at zero
    for (int d = 0; d < 3; ++d) {
        s->dU[d].assign(n, 0.0); s->dV[d].assign(n, 0.0); s->dW[d].assign(n, 0.0); s->dT[d].assign(n, 0.0);
        for (int e = 0; e < 3; ++e) s->grad[d][e].assign(n, 0.0);
    }
    for (int e = 0; e < 15; ++e) { s->pre[e].assign(n, 0.0); s->div[e].assign(n, 0.0); }
    if (cfg->les == ORC_DSM) {
        Vec *d[] = {&s->S00, &s->S01, &s->S02, &s->S11, &s->S12, &s->S22, &s->Smag, &s->rho_hat,
                    &s->Smag_hat, &s->rhoU_hat, &s->rhoV_hat, &s->rhoW_hat};
        for (Vec *v : d) v->assign(n, 0.0);
        for (int k = 0; k < 19; ++k) s->work[k].assign(n, 0.0);
    }
    // setInitialConditions, src/CurvilinearCSolver_BC.cpp:30-63
    for (long ip = 0; ip < n; ++ip) {
        s->U[ip] = U0[ip]; s->V[ip] = V0[ip]; s->W[ip] = W0[ip]; s->p[ip] = p0[ip];
        s->rho1[ip] = rho0[ip];
        s->rhoU1[ip] = s->rho1[ip] * s->U[ip];
        s->rhoV1[ip] = s->rho1[ip] * s->V[ip];
        s->rhoW1[ip] = s->rho1[ip] * s->W[ip];
        s->Ucurv[ip] = s->J11[ip] * s->U[ip] + s->J12[ip] * s->V[ip] + s->J13[ip] * s->W[ip];
        s->Vcurv[ip] = s->J21[ip] * s->U[ip] + s->J22[ip] * s->V[ip] + s->J23[ip] * s->W[ip];
        s->Wcurv[ip] = s->J31[ip] * s->U[ip] + s->J32[ip] * s->V[ip] + s->J33[ip] * s->W[ip];
    }
    for (long ip = 0; ip < n; ++ip) {
        s->rhoE1[ip] = s->ig.solverhoE(s->rho1[ip], s->p[ip], s->U[ip], s->V[ip], s->W[ip]);
        s->T[ip] = s->ig.solveT(s->rho1[ip], s->p[ip]);
        s->mu[ip] = s->ig.solveMu(s->T[ip]);
        s->sos[ip] = s->ig.solveSOS(s->rho1[ip], s->p[ip]);
    }
    s->wallIC();
    s->timeStep = cfg->timestep0;
    s->time = cfg->time0;
    s->dt = cfg->dt;
    s->cfl = cfg->cfl;
    // src/Stats.hpp:42-250: zero fields and weights, prev_time = the current time
    if (cfg->velocity_stats)
        for (auto &v : s->vstat) v.assign(n, 0.0);
    if (cfg->thermo_stats)
        for (auto &v : s->tstat) v.assign(n, 0.0);
    s->statsPrevTime = cfg->time0;
    return s;
}

int orc_solver_set_sponge(orc_solver *s, const double *sigma, const double *const avg[5]) {
    s->spongeFlag = true;
    s->sigma.assign(sigma, sigma + s->n);
    const Vec *q[5] = {&s->rho1, &s->rhoU1, &s->rhoV1, &s->rhoW1, &s->rhoE1};
    for (int e = 0; e < 5; ++e) {
        if (avg) s->spgAvg[e].assign(avg[e], avg[e] + s->n);
        else s->spgAvg[e] = *q[e];
    }
    return 0;
}

// src/SpongeBC.cpp:60-260: sigma = max over sponge faces of
// strength (0.068 s^2 + 0.845 s^8), s the normalised depth into the layer
int orc_sponge_sigma(int Nx, int Ny, int Nz, const double *x, const double *y, const double *z,
                     const int faces[6], int kind, double strength, const double perc[6], int cyl_orient,
                     const double axis[3], double rmin, double *sigma) {
    const long n = (long)Nx * Ny * Nz;
    const double *X[3] = {x, y, z};
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {  // msh->x_min / x_max
        lo[d] = *std::min_element(X[d], X[d] + n);
        hi[d] = *std::max_element(X[d], X[d] + n);
    }
    for (long ip = 0; ip < n; ++ip) sigma[ip] = 0.0;
    auto ramp = [&](double s) { return strength * (0.068 * std::pow(s, 2.0) + 0.845 * std::pow(s, 8.0)); };
    if (kind == 0) {
        for (int f = 0; f < 6; ++f) {
            if (faces[f] != ORC_FACE_SPONGE) continue;
            const int d = f / 2;
            const double L = perc[f] * (hi[d] - lo[d]);
            for (long ip = 0; ip < n; ++ip) {
                const double v = X[d][ip];
                if (f % 2 == 0) {
                    const double dd = v - lo[d];
                    if (dd < L) sigma[ip] = std::fmax(ramp((L - dd) / L), sigma[ip]);
                } else {
                    const double dd = hi[d] - v;
                    if (dd < L) sigma[ip] = std::fmax(ramp((v - (hi[d] - L)) / L), sigma[ip]);
                }
            }
        }
        return 0;
    }
    if (cyl_orient != 2) return -1;
    const double rx = std::fmax(std::fabs(lo[0] - axis[0]), std::fabs(hi[0] - axis[0]));
    const double ry = std::fmax(std::fabs(lo[1] - axis[1]), std::fabs(hi[1] - axis[1]));
    const double rmax = std::fmax(rx, ry);
    if (faces[3] == ORC_FACE_SPONGE) {  // only the Y1 face, :205-219
        for (long ip = 0; ip < n; ++ip) {
            const double r = std::sqrt(std::pow(x[ip] - axis[0], 2.0) + std::pow(y[ip] - axis[1], 2.0));
            if (r > rmin) sigma[ip] = std::fmax(ramp((r - rmin) / (rmax - rmin)), sigma[ip]);
        }
    }
    return 0;
}

int orc_solver_steps(orc_solver *s, int n) {
    for (int i = 0; i < n; ++i) s->step();
    return 0;
}

const double *orc_solver_field(const orc_solver *s, int id) {
    switch (id) {
        case ORC_F_RHO: return s->rho1.data();
        case ORC_F_RHOU: return s->rhoU1.data();
        case ORC_F_RHOV: return s->rhoV1.data();
        case ORC_F_RHOW: return s->rhoW1.data();
        case ORC_F_RHOE: return s->rhoE1.data();
        case ORC_F_U: return s->U.data();
        case ORC_F_V: return s->V.data();
        case ORC_F_W: return s->W.data();
        case ORC_F_P: return s->p.data();
        case ORC_F_T: return s->T.data();
        case ORC_F_MU: return s->mu.data();
        case ORC_F_SOS: return s->sos.data();
        case ORC_F_MUSGS: return s->mu_sgs.data();
        case ORC_F_TAUKK: return s->taukk.data();
        case ORC_F_KSGS: return s->k_sgs.data();
        case ORC_F_C: return s->C.data();
        case ORC_F_CI: return s->CI.data();
        case ORC_F_PRT: return s->Prt.data();
        case ORC_F_RHOK: return s->rhok.data();
        case ORC_F_RHOUK: return s->rhoUk.data();
        case ORC_F_RHOVK: return s->rhoVk.data();
        case ORC_F_RHOWK: return s->rhoWk.data();
        case ORC_F_RHOEK: return s->rhoEk.data();
        case ORC_F_RHOK2: return s->rhok2.data();
        case ORC_F_RHOUK2: return s->rhoUk2.data();
        case ORC_F_RHOVK2: return s->rhoVk2.data();
        case ORC_F_RHOWK2: return s->rhoWk2.data();
        case ORC_F_RHOEK2: return s->rhoEk2.data();
        case ORC_F_RHO2: return s->rho2.data();
        default:
            if (id >= ORC_F_UAVG && id <= ORC_F_VWREY)
                return s->c.velocity_stats ? s->vstat[id - ORC_F_UAVG].data() : nullptr;
            if (id >= ORC_F_RHOAVG && id <= ORC_F_TRMS)
                return s->c.thermo_stats ? s->tstat[id - ORC_F_RHOAVG].data() : nullptr;
            return nullptr;
    }
}

int orc_solver_rhs(orc_solver *s, double *const out[5]) {
    s->rkStep = 1;
    s->rkLast = false;
    s->preStepDerivatives();
    static const int idx[5][3] = {{0, 5, 10}, {1, 6, 11}, {2, 7, 12}, {3, 8, 13}, {4, 9, 14}};
    for (int q = 0; q < 5; ++q)
        for (long ip = 0; ip < s->n; ++ip)
            out[q][ip] = s->J[ip] * (-s->div[idx[q][0]][ip] - s->div[idx[q][1]][ip] - s->div[idx[q][2]][ip]);
    return 0;
}

double orc_rk_amplification(int rk, int dual_step, double z) {
    orc_solver s;
    s.c = orc_config{};
    s.c.rk = rk;
    s.n = 1;
    s.dualStep = dual_step;
    Vec *regs[] = {&s.rho1, &s.rhoU1, &s.rhoV1, &s.rhoW1, &s.rhoE1, &s.rhok, &s.rhoUk, &s.rhoVk,
                   &s.rhoWk, &s.rhoEk, &s.rhok2, &s.rhoUk2, &s.rhoVk2, &s.rhoWk2, &s.rhoEk2,
                   &s.rho2, &s.rhoU2, &s.rhoV2, &s.rhoW2, &s.rhoE2};
    for (Vec *v : regs) v->assign(1, 0.0);
    s.rho1[0] = 1.0;
    const int S = s.n_stages();
    for (int st = 1; st <= S; ++st) {
        s.rkStep = st;
        const double y = st == 1 ? s.rho1[0] : s.rhok[0];  // stage state: rho1, then rhok
        s.rhok2[0] = z * y;
        s.updateConservedData();
    }
    return s.rho2[0];
}

double orc_solver_time(const orc_solver *s) { return s->time; }
double orc_solver_stats_weight(const orc_solver *s, int which) { return which ? s->thermoWeight : s->velWeight; }
double orc_solver_dt(const orc_solver *s) { return s->dt; }
int orc_solver_timestep(const orc_solver *s) { return s->timeStep; }
void orc_solver_destroy(orc_solver *s) { delete s; }

}  // extern "C"
