// Fast line kernels of (W, H) = (2, 3): Penta10.
#include "lineop_fast_kernels.h"

namespace mc3d {

hipError_t lineop_launch_fast_w2h3(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return launch_fast_mm<2, 3>(op, a, nfields, st);
}

}  // namespace mc3d
