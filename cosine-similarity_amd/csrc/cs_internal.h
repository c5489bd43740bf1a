// cs_internal.h -- helpers shared by the translation units of libcsamd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cosine_similarity_amd.h"

namespace cs {

// Record a formatted error for cs_last_error() and return `code`.
cs_status fail(cs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves aligned sub-buffers out of one caller-provided workspace.  With a
// null base it only measures.
struct Carver {
    char *base;
    size_t used = 0;
    explicit Carver(void *b) : base(static_cast<char *>(b)) {}
    template <class T>
    T *take(size_t count) {
        used = align_up(used, 256);
        T *p = base ? reinterpret_cast<T *>(base + used) : nullptr;
        used += count * sizeof(T);
        return p;
    }
};

}  // namespace cs

#define CS_HIP_TRY(expr)                                                                   \
    do {                                                                                   \
        hipError_t cs_e_ = (expr);                                                         \
        if (cs_e_ != hipSuccess)                                                           \
            return ::cs::fail(CS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(cs_e_)); \
    } while (0)

#define CS_LAUNCH_CHECK(name)                                                              \
    do {                                                                                   \
        hipError_t cs_e_ = hipGetLastError();                                              \
        if (cs_e_ != hipSuccess)                                                           \
            return ::cs::fail(CS_ERR_HIP, "launch of %s failed: %s", name,                  \
                              hipGetErrorString(cs_e_));                                   \
    } while (0)
