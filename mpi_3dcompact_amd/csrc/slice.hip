// Index-plane gather: one thread per output point, stores coalesced along a1.
#include "slice.h"

namespace mc3d {

__global__ void __launch_bounds__(256) k_slice(const double *__restrict__ f, double *__restrict__ out, int dir,
                                               int index, int Nx, int Ny, int Nz) {
    const int n1 = dir == 0 ? Ny : dir == 1 ? Nz : Nx;
    const int n2 = dir == 0 ? Nz : dir == 1 ? Nx : Ny;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)n1 * n2) return;
    const int a1 = (int)(q % n1), a2 = (int)(q / n1);
    long p;
    if (dir == 0)
        p = ((long)a2 * Ny + a1) * Nx + index;
    else if (dir == 1)
        p = ((long)a1 * Ny + index) * Nx + a2;
    else
        p = ((long)index * Ny + a2) * Nx + a1;
    out[q] = f[p];
}

hipError_t launch_slice(const double *f, double *out, int dir, int index, int Nx, int Ny, int Nz,
                        hipStream_t st) {
    const long n = dir == 0 ? (long)Ny * Nz : dir == 1 ? (long)Nz * Nx : (long)Nx * Ny;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_slice, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, out, dir, index, Nx, Ny,
                       Nz);
    return hipGetLastError();
}

}  // namespace mc3d
