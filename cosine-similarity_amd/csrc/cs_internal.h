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

// x from lane (lane ^ J) of the wave.  J <= 8 stays within a 16-lane row and
// uses DPP lane moves (VALU, no LDS round trip): quad_perm [1,0,3,2] and
// [2,3,0,1] for 1 and 2, row_half_mirror then quad_perm [3,2,1,0] for 4 (lane
// i of a half-row reads 7 - i, then quad lane k reads 3 - k: i ^ 4),
// row_ror:8 for 8.  J = 16 is a ds_swizzle xor within 32 lanes, J = 32 a
// bpermute.
template <int J>
__device__ __forceinline__ int xor_lane(int x) {
    if constexpr (J == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    if constexpr (J == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    if constexpr (J == 4)
        return __builtin_amdgcn_update_dpp(
            0, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
    if constexpr (J == 8) return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);
    if constexpr (J == 16) return __builtin_amdgcn_ds_swizzle(x, 0x401F);
    return __shfl_xor(x, 32);
}

template <int J>
__device__ __forceinline__ double xor_lane_double(double x) {
    const int lo = xor_lane<J>(__double2loint(x));
    const int hi = xor_lane<J>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

template <int J>
__device__ __forceinline__ float xor_lane_float(float x) {
    return __int_as_float(xor_lane<J>(__float_as_int(x)));
}

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
