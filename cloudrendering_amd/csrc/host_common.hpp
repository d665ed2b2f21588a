// Host-side helpers shared by the library's C-ABI translation units: the last-error slot, HIP status checks and
// a growable device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/crvpt.h"

namespace crvpt {

// Records msg as the message crvpt_last_error returns and returns code.
int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) return ::crvpt::fail(CRVPT_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Device allocation that grows on demand (never shrinks); release() frees it.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && ptr) return hipSuccess;
        // A kernel on any stream may still read the old allocation.
        if (ptr) (void)hipDeviceSynchronize();
        release();
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            bytes = 0;
            return e;
        }
        bytes = n;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(ptr);
    }
};

}  // namespace crvpt
