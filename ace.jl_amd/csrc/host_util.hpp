// Host helpers shared by the translation units of libacehip.so: the
// thread-local error message behind ace_last_error(), the HIP_TRY macro and
// a grow-only device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

// last error message of the calling thread (one instance per process)
inline thread_local std::string g_err;

inline int set_err(const std::string& s) {
    g_err = s;
    return -1;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return set_err(std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t count) {
        if (count <= n) return 0;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            n = 0;
        }
        HIP_TRY(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
        return 0;
    }
    int upload(const T* h, size_t count) {
        if (ensure(count)) return -1;
        if (count) HIP_TRY(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }
};
