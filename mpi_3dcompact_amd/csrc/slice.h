// Index-plane gather of one field for image and slice output.
#pragma once
#include <hip/hip_runtime.h>

namespace mc3d {

// Plane `index` normal to `dir` of an x-fastest (Nx, Ny, Nz) field, written as
// out[a2 * n1 + a1] with the in-plane axes of the reference's image planes
// (src/CurvilinearCSolver_Core.cpp:1494-1650): dir 0 -> (a1, a2) = (j, k),
// dir 1 -> (k, i), dir 2 -> (i, j).
hipError_t launch_slice(const double *f, double *out, int dir, int index, int Nx, int Ny, int Nz,
                        hipStream_t st);

}  // namespace mc3d
