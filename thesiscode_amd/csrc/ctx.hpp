// Library context shared by the C ABI sources (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "scrimp_hip.h"

namespace mp {

constexpr int64_t kMaxProfile = int64_t(1) << 30;

// Grow-only device arena owned by a context; sub-allocations are 256-byte aligned.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    int reserve(size_t bytes) {
        if (bytes <= cap) return MP_OK;
        if (base) (void)hipFree(base);
        base = nullptr;
        cap = 0;
        if (hipMalloc(&base, bytes) != hipSuccess) {
            base = nullptr;
            return MP_ERR_OOM;
        }
        cap = bytes;
        return MP_OK;
    }
    template <class T>
    T* take(size_t count) {
        const size_t b = (count * sizeof(T) + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += b;
        return p;
    }
};

inline size_t aligned_bytes(size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); }

}  // namespace mp

struct mp_ctx {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    mp::Arena work;  // per-call scratch
    mp::Arena io;    // staging for host-pointer calls
    std::mutex lock;
};
