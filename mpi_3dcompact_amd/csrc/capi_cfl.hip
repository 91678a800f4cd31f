// C ABI of the CFL time step over caller-owned fields (include/mc3d.h):
// CurvilinearCSolver::calcDtFromCFL (src/CurvilinearCSolver_Core.cpp:189-237)
// without the solver around it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_common.h"
#include "mc3d.h"
#include "pointwise.h"

namespace {

int check_args(const double *const fields[7], long npts, int members, const double spacing[3], double cfl,
               const double *out) {
    if (!fields || !spacing || !out) return mc3d::fail(MC3D_EINVAL, "null argument");
    for (int k = 0; k < 7; ++k)
        if (!fields[k]) return mc3d::fail(MC3D_EINVAL, "null field");
    if (npts < 1 || members < 1) return mc3d::fail(MC3D_EINVAL, "npts and members must be positive");
    if (npts * members > (long)2147483647) return mc3d::fail(MC3D_EINVAL, "more than 2^31 points");
    for (int d = 0; d < 3; ++d)
        if (!(spacing[d] > 0.0)) return mc3d::fail(MC3D_EINVAL, "spacings must be positive");
    if (!(cfl > 0.0)) return mc3d::fail(MC3D_EINVAL, "cfl must be positive");
    return MC3D_OK;
}

// partial maxima per member into partial[member * cfl_blocks(npts) + block],
// then dt and the maximum per member into out
hipError_t launch(const double *const fields[7], long npts, int members, const double spacing[3], double cfl,
                  double *partial, double *out, hipStream_t st) {
    mc3d::Prim pr{};
    pr.U = const_cast<double *>(fields[0]);
    pr.V = const_cast<double *>(fields[1]);
    pr.W = const_cast<double *>(fields[2]);
    pr.sos = const_cast<double *>(fields[3]);
    mc3d::Metrics m{};
    m.J11 = fields[4];
    m.J22 = fields[5];
    m.J33 = fields[6];
    hipError_t e = mc3d::launch_cfl(pr, m, spacing[0], spacing[1], spacing[2], partial, npts, members, st);
    if (e != hipSuccess) return e;
    return mc3d::launch_cfl_out(partial, mc3d::cfl_blocks(npts), members, cfl, out, st);
}

// Partial maxima of mc3d_cfl_dt_noalloc: cfl_blocks(npts) <= 2048 doubles per
// box for at most kScratchBoxes boxes (1 MiB), held by every GPU the library
// runs on until the process ends.  Host code never passes `scratch` itself to
// a kernel (on the host it names a stand-in, not device memory): it passes the
// device address read once per device with hipGetSymbolAddress.
constexpr int kScratchBoxes = 64;
constexpr long kScratchDoubles = 2048L * kScratchBoxes;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up on the first call there
int scratch_address(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return mc3d::fail(MC3D_EHIP, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return mc3d::fail(MC3D_EINVAL, "device index beyond 63");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mc3d::fail(MC3D_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = scratch_ptr[dev];
    return MC3D_OK;
}

}  // namespace

extern "C" {

int mc3d_cfl_dt(const double *const fields[7], long npts, int members, const double spacing[3], double cfl,
                double *out, void *stream) {
    int rc = check_args(fields, npts, members, spacing, cfl, out);
    if (rc) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    double *partial = nullptr;
    const size_t bytes = sizeof(double) * (size_t)mc3d::cfl_blocks(npts) * members;
    if (hipMalloc(&partial, bytes) != hipSuccess) return mc3d::fail(MC3D_ENOMEM, "hipMalloc of CFL partials failed");
    hipError_t e = launch(fields, npts, members, spacing, cfl, partial, out, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipFree(partial);
    if (e != hipSuccess) return mc3d::fail(MC3D_EHIP, std::string("CFL launch: ") + hipGetErrorString(e));
    if (es != hipSuccess) return mc3d::fail(MC3D_EHIP, std::string("CFL: ") + hipGetErrorString(es));
    return MC3D_OK;
}

int mc3d_cfl_dt_noalloc(const double *const fields[7], long npts, int members, const double spacing[3],
                        double cfl, double *out, void *stream) {
    int rc = check_args(fields, npts, members, spacing, cfl, out);
    if (rc) return rc;
    if ((long)mc3d::cfl_blocks(npts) * members > kScratchDoubles)
        return mc3d::fail(MC3D_EINVAL, "mc3d_cfl_dt_noalloc takes at most 64 boxes (its scratch holds 2048 "
                                       "partial maxima per box); use mc3d_cfl_dt");
    double *partial = nullptr;
    rc = scratch_address(&partial);
    if (rc) return rc;
    // every workgroup of k_cfl writes its own partial, so nothing is cleared
    const hipError_t e = launch(fields, npts, members, spacing, cfl, partial, out, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return mc3d::fail(MC3D_EHIP, std::string("CFL launch: ") + hipGetErrorString(e));
    return MC3D_OK;
}

}  // extern "C"
