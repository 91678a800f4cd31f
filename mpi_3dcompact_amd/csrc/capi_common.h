// Error plumbing shared by the C ABI translation units.
#pragma once

#include <string>

namespace mc3d {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string &msg);

}  // namespace mc3d
