// Fast line kernels of (W, H) = (0, 3): CommN5Exp.
#include "lineop_fast_kernels.h"

namespace mc3d {

hipError_t lineop_launch_fast_w0h3(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return launch_fast_mm<0, 3>(op, a, nfields, st);
}

}  // namespace mc3d
