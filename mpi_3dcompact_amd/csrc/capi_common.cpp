// Thread-local error text and device queries of the C ABI.
#include <hip/hip_runtime.h>

#include <string>

#include "capi_common.h"
#include "mc3d.h"

namespace mc3d {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace mc3d

extern "C" {

const char *mc3d_last_error(void) { return mc3d::g_last_error.c_str(); }

int mc3d_device_count(int *count) {
    if (!count) return mc3d::fail(MC3D_EINVAL, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return mc3d::fail(MC3D_EHIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return MC3D_OK;
}

}  // extern "C"
