// Fast line kernels of (W, H) = (0, 1): CD2.
#include "lineop_fast_kernels.h"

namespace mc3d {

hipError_t lineop_launch_fast_w0h1(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return launch_fast_mm<0, 1>(op, a, nfields, st);
}

}  // namespace mc3d
