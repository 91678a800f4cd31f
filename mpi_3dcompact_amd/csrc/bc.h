// Boundary handling of the solver step: walls and sponge
// (src/CurvilinearCSolver_BC.cpp:3-842, src/SpongeBC.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "pointwise.h"

namespace mc3d {

// One wall face: velocities zero, momentum zero (rho * 0 for a moving wall,
// whose velocities are zero in the reference), the temperature optionally
// replaced by the zero-slope extrapolation of the six points inward along
// the face normal (calcNeumann, src/Pade6.cpp:395-397, src/Penta10.cpp:483-485,
// src/CD2.cpp:183-185).  With stage set, p = rho R T and rhoE = p/(gamma-1) +
// rho |u|^2 / 2 follow from that state (preStepBCHandling, :372-700).
struct WallArgs {
    int face;
    int Nx, Ny, Nz;
    int neumann;  // replace T
    int scheme;   // derivative scheme of the face's direction (MC3D_CD2 / PADE6 / PENTA10)
    int moving;
    int stage;
    Cons q;       // state whose momentum (and rhoE) the fix writes
    Prim pr;
    Gas g;
};
hipError_t launch_wall_face(const WallArgs &a, hipStream_t st);
// after the initial wall fixes: p = rho R T, sos, rhoE everywhere (:311-322)
hipError_t launch_wall_ic_update(Cons r1, Prim pr, Gas g, long n, hipStream_t st);

}  // namespace mc3d
