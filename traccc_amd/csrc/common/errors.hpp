// Error handling of the C ABI: library code throws ckf_error (or a HIP error
// through hip_check); every extern "C" entry point runs its body in
// guarded(), which turns the exception into a ckf_status and keeps the
// message for ckf_last_error().  No exception crosses the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "traccc_amd/ckf.h"

namespace tamd {

// the calling thread's last error message (defined in common/last_error.cpp)
std::string& last_error();

struct ckf_error : std::runtime_error {
    ckf_status code;
    ckf_error(ckf_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw ckf_error(e == hipErrorOutOfMemory ? CKF_ERR_OUT_OF_MEMORY : CKF_ERR_DEVICE,
                        std::string(what) + ": " + hipGetErrorString(e));
    }
}

template <typename F>
ckf_status guarded(F&& f) {
    try {
        f();
        last_error().clear();
        return CKF_OK;
    } catch (const ckf_error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return CKF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return CKF_ERR_INTERNAL;
    }
}

}  // namespace tamd
