// Metric terms of a single-block curvilinear mesh: the part of
// AbstractSingleBlockMesh that turns coordinates into J and J_ij
// (src/AbstractSingleBlockMesh.cpp:23-1309, translations :2019-2085).
#pragma once

#include <hip/hip_runtime.h>

#include "solver.hpp"

namespace mc3d {

struct MeshSpec {
    int N[3];
    bool periodic[3];
    bool bw7[3];     // 7-point explicit stencil (Penta10): the wrap slot Nm3 exists
    double T[3][3];  // T[d][e]: component e of the translation across periodic direction d
};

// out[10] = {J, J11, J12, J13, J21, J22, J23, J31, J32, J33}; x, y, z and out
// are device fields of N[0]*N[1]*N[2] points, x fastest.  dv[d] differentiates
// along direction d.  Synchronous on st; allocates its scratch for the call.
void solve_for_jacobians(AbstractDerivatives *const dv[3], const MeshSpec &m, const double *x,
                         const double *y, const double *z, double *const out[10], hipStream_t st);

}  // namespace mc3d
