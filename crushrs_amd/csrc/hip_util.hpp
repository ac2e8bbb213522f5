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

// A fixed device region that buffers are carved from instead of hipMalloc:
// 256-byte aligned pieces in order, never freed one by one.  While an
// ArenaScope is alive on a thread, every DevBuf that thread allocates comes
// from its arena.
struct DevArena {
    uint8_t* base = nullptr;  // device address of the region
    size_t cap = 0, used = 0;
    const char* owner = "";   // named in the error when the region is full
    void* take(size_t bytes) {
        const size_t off = (used + 255) & ~size_t(255);
        if (off + bytes > cap)
            fail(CRUSH_E_INPUT, "%s: the model needs more than the %zu bytes of its device scratch", owner, cap);
        used = off + bytes;
        return base + off;
    }
};

inline DevArena*& current_arena() {
    static thread_local DevArena* a = nullptr;
    return a;
}

struct ArenaScope {
    DevArena* prev;
    explicit ArenaScope(DevArena* a) : prev(current_arena()) { current_arena() = a; }
    ~ArenaScope() { current_arena() = prev; }
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = false;  // from hipMalloc (false: from an arena, or none)
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = false;
    }
    void alloc(size_t count) {
        release();
        n = std::max<size_t>(count, 1);
        if (DevArena* a = current_arena()) {
            p = static_cast<T*>(a->take(n * sizeof(T)));
            return;
        }
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            fail(CRUSH_E_OOM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
        }
        owned = true;
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
