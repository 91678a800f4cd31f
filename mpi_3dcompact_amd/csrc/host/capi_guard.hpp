// Exception barrier of the host-class C ABI entry points: every mc3d::Error,
// allocation failure or other exception becomes an error code plus the
// thread's last-error text.
#pragma once

#include <exception>
#include <new>

#include "../capi_common.h"
#include "mc3d.h"
#include "solver.hpp"

namespace mc3d {

template <class F>
int guarded(F &&f) {
    try {
        f();
        return MC3D_OK;
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(MC3D_ENOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(MC3D_EINVAL, e.what());
    }
}

}  // namespace mc3d
