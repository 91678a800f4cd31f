// Boundary handling of the solver: walls (src/CurvilinearCSolver_BC.cpp:3-842)
// and the sponge (src/SpongeBC.cpp).  All wall velocities are zero, as the
// reference initialises them (src/CurvilinearCSolver.hpp:219-221); the sponge
// averages stay at their initial values (updateSponge is not called by the
// reference's postStep, src/CurvilinearCSolver.hpp:359).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "../bc.h"
#include "solver.hpp"

#include <cstdio>
#include <fstream>
#include <vector>

namespace mc3d {

namespace {

bool is_wall(Options::BCKind k) {
    return k == Options::ADIABATIC_WALL || k == Options::CONST_T_WALL || k == Options::MOVING_ADIABATIC_WALL;
}

}  // namespace

void CurvilinearCSolver::setupBoundaries() {
    const Options::BCKind face[6] = {opt.bcX0, opt.bcX1, opt.bcY0, opt.bcY1, opt.bcZ0, opt.bcZ1};
    for (int f = 0; f < 6; ++f) {
        role_[f] = is_wall(face[f]) ? kFaceWall : face[f] == Options::SPONGE ? kFaceSponge : kFaceNone;
        walls_ = walls_ || role_[f] == kFaceWall;
        sponge_ = sponge_ || role_[f] == kFaceSponge;
    }
    for (int d = 0; d < 3; ++d)
        if (role_[2 * d] == kFaceWall || role_[2 * d + 1] == kFaceWall) cont_[d] = alloc_field();
    if (sponge_) {
        sigma_ = alloc_field();
        for (auto &a : spgAvg_) a = alloc_field();
    }
}

// src/CurvilinearCSolver_BC.cpp:63-322: static walls X0..Z1, then moving walls;
// the Neumann temperature on ADIABATIC_WALL faces, on X1 for CONST_T_WALL
// instead (:103, as the reference has it) and on every moving wall; then, with
// any wall, p, sos and rhoE from T everywhere.
void CurvilinearCSolver::wallInitialConditions() {
    if (!walls_) return;
    const Options::BCKind face[6] = {opt.bcX0, opt.bcX1, opt.bcY0, opt.bcY1, opt.bcZ0, opt.bcZ1};
    const Options::FDType ft[3] = {opt.xFDType, opt.yFDType, opt.zFDType};
    WallArgs a{};
    a.Nx = dom.gNx; a.Ny = dom.gNy; a.Nz = dom.gNz;
    a.q = r1; a.pr = pr; a.g = ig;
    a.stage = 0;
    for (int f = 0; f < 6; ++f) {
        if (face[f] != Options::ADIABATIC_WALL && face[f] != Options::CONST_T_WALL) continue;
        a.face = f;
        a.scheme = (int)ft[f / 2];
        a.moving = 0;
        a.neumann = f == 1 ? face[f] == Options::CONST_T_WALL : face[f] == Options::ADIABATIC_WALL;
        hip_check(launch_wall_face(a, st), "wall initial condition");
    }
    for (int f = 0; f < 6; ++f) {
        if (face[f] != Options::MOVING_ADIABATIC_WALL) continue;
        a.face = f;
        a.scheme = (int)ft[f / 2];
        a.moving = 1;
        a.neumann = 1;
        hip_check(launch_wall_face(a, st), "wall initial condition");
    }
    hip_check(launch_wall_ic_update(r1, pr, ig, N, st), "wall initial condition");
}

// src/CurvilinearCSolver_BC.cpp:372-700, at the start of every stage on the
// stage's state: static walls X0..Z1 (Neumann temperature on ADIABATIC_WALL),
// then moving walls (always Neumann); p and rhoE follow from T.
void CurvilinearCSolver::preStepBCHandling() {
    if (!walls_) return;
    const Options::BCKind face[6] = {opt.bcX0, opt.bcX1, opt.bcY0, opt.bcY1, opt.bcZ0, opt.bcZ1};
    const Options::FDType ft[3] = {opt.xFDType, opt.yFDType, opt.zFDType};
    WallArgs a{};
    a.Nx = dom.gNx; a.Ny = dom.gNy; a.Nz = dom.gNz;
    a.q = state(); a.pr = pr; a.g = ig;
    a.stage = 1;
    for (int pass = 0; pass < 2; ++pass)
        for (int f = 0; f < 6; ++f) {
            const bool moving = face[f] == Options::MOVING_ADIABATIC_WALL;
            const bool fixed = face[f] == Options::ADIABATIC_WALL || face[f] == Options::CONST_T_WALL;
            if (pass == 0 ? !fixed : !moving) continue;
            a.face = f;
            a.scheme = (int)ft[f / 2];
            a.moving = moving;
            a.neumann = moving || face[f] == Options::ADIABATIC_WALL;
            timed("k_wall_face", 16.0 / (double)(f < 2 ? dom.gNx : f < 4 ? dom.gNy : dom.gNz),
                  [&] { hip_check(launch_wall_face(a, st), "wall boundary"); });
        }
}

RkBc CurvilinearCSolver::rkBoundaryArgs() const {
    RkBc b{};
    b.active = walls_ || sponge_;
    for (int f = 0; f < 6; ++f) b.role[f] = role_[f];
    b.Nx = dom.gNx; b.Ny = dom.gNy; b.Nz = dom.gNz;
    for (int d = 0; d < 3; ++d) b.cont[d] = cont_[d];
    b.sponge = sponge_ ? 1 : 0;
    b.sigma = sigma_;
    for (int q = 0; q < 5; ++q) b.avg[q] = spgAvg_[q];
    b.qs = state();
    return b;
}

// src/CurvilinearCSolver_BC.cpp:324-345: the averages start as the initial
// conserved fields (after the wall fixes), or, with SPONGEAVGFROMRESTART, as the
// five fields of SPONGEAVGFILENAME (SpongeBC::readSpongeAvgFromRestart,
// src/SpongeBC.cpp:426-475: rho, rhoU, rhoV, rhoW, rhoE, x fastest, no header)
void CurvilinearCSolver::initSpongeAverages() {
    if (!sponge_ || avg_given_) return;
    if (opt.spongeFromRestart) {
        std::ifstream in(opt.sponge_filename, std::ios::binary);
        if (!in) throw Error(MC3D_EINVAL, "cannot open sponge averages " + opt.sponge_filename);
        std::vector<double> buf((size_t)N);
        for (int q = 0; q < 5; ++q) {
            in.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)(sizeof(double) * N));
            if (!in) throw Error(MC3D_EINVAL, "sponge averages file too short: " + opt.sponge_filename);
            hip_check(hipMemcpyAsync(spgAvg_[q], buf.data(), sizeof(double) * N, hipMemcpyHostToDevice, st),
                      "sponge averages");
            hip_check(hipStreamSynchronize(st), "sponge averages");
        }
        std::printf(" > Initializing sponge average field from restart file: %s\n", opt.sponge_filename.c_str());
        return;
    }
    for (int q = 0; q < 5; ++q)
        hip_check(hipMemcpyAsync(spgAvg_[q], r1.q[q], sizeof(double) * N, hipMemcpyDeviceToDevice, st),
                  "sponge averages");
}

void CurvilinearCSolver::setSponge(const double *sigma_host, const double *const avg_host[5]) {
    if (!sponge_) throw Error(MC3D_ECONFIG, "no face is a SPONGE");
    hip_check(hipMemcpyAsync(sigma_, sigma_host, sizeof(double) * N, hipMemcpyHostToDevice, st), "sponge");
    if (avg_host) {
        for (int q = 0; q < 5; ++q)
            hip_check(hipMemcpyAsync(spgAvg_[q], avg_host[q], sizeof(double) * N, hipMemcpyHostToDevice, st),
                      "sponge averages");
        avg_given_ = true;
    }
    hip_check(hipStreamSynchronize(st), "sponge");
    sigma_set_ = true;
}

// SpongeBC::initRectSpongeBC / initCylSpongeBC (src/SpongeBC.cpp:60-260):
// sigma = max over the sponge faces of strength (0.068 s^2 + 0.845 s^8), s the
// normalised depth into the layer, over the mesh's coordinate ranges.  The
// cylindrical layer exists for axis orientation 2 on the Y1 face only (:205-219).
void CurvilinearCSolver::setSpongeFromMesh(const double *x, const double *y, const double *z) {
    if (!sponge_) return;
    const double *X[3] = {x, y, z};
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = *std::min_element(X[d], X[d] + N);
        hi[d] = *std::max_element(X[d], X[d] + N);
    }
    std::vector<double> sg((size_t)N, 0.0);
    const double S = opt.spongeStrength;
    auto ramp = [S](double s) { return S * (0.068 * std::pow(s, 2.0) + 0.845 * std::pow(s, 8.0)); };
    if (opt.spongeKind == Options::RECTILINEAR) {
        for (int f = 0; f < 6; ++f) {
            if (role_[f] != kFaceSponge) continue;
            const int d = f / 2;
            const double L = opt.spongeRectPerc[f] * (hi[d] - lo[d]);
            for (long p = 0; p < N; ++p) {
                const double v = X[d][p];
                if (f % 2 == 0) {
                    const double dd = v - lo[d];
                    if (dd < L) sg[(size_t)p] = std::fmax(ramp((L - dd) / L), sg[(size_t)p]);
                } else {
                    const double dd = hi[d] - v;
                    if (dd < L) sg[(size_t)p] = std::fmax(ramp((v - (hi[d] - L)) / L), sg[(size_t)p]);
                }
            }
        }
    } else {
        if (opt.spongeCylAxisOrient != 2)
            throw Error(MC3D_ECONFIG, "the cylindrical sponge exists for SPONGE.CYLAXIS_ORIENT = 2 only "
                                      "(src/SpongeBC.cpp:222-225)");
        const double *ax = opt.spongeCylAxis;
        const double rx = std::fmax(std::fabs(lo[0] - ax[0]), std::fabs(hi[0] - ax[0]));
        const double ry = std::fmax(std::fabs(lo[1] - ax[1]), std::fabs(hi[1] - ax[1]));
        const double rmax = std::fmax(rx, ry);
        if (role_[3] == kFaceSponge)
            for (long p = 0; p < N; ++p) {
                const double r = std::sqrt(std::pow(x[p] - ax[0], 2.0) + std::pow(y[p] - ax[1], 2.0));
                if (r > opt.rMin) sg[(size_t)p] = std::fmax(ramp((r - opt.rMin) / (rmax - opt.rMin)), sg[(size_t)p]);
            }
    }
    setSponge(sg.data(), nullptr);
}

}  // namespace mc3d
