// Fast line kernels of (W, H) = (2, 5): CommN5Pade.
#include "lineop_fast_kernels.h"

namespace mc3d {

hipError_t lineop_launch_fast_w2h5(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return launch_fast_mm<2, 5>(op, a, nfields, st);
}

}  // namespace mc3d
