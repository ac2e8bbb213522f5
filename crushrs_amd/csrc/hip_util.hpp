// HIP runtime helpers shared by the host-side solver drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace crush {

#define HIP_OK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) ::crush::fail(CRUSH_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        n = std::max<size_t>(count, 1);
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            fail(CRUSH_E_OOM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
        }
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) { HIP_OK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
};

inline void require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(CRUSH_E_DEVICE, "no HIP device available");
}

}  // namespace crush
