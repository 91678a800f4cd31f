// Fast line kernels of (W, H) = (1, 5): Compact10.
#include "lineop_fast_kernels.h"

namespace mc3d {

hipError_t lineop_launch_fast_w1h5(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return launch_fast_mm<1, 5>(op, a, nfields, st);
}

}  // namespace mc3d
