// records.hip -- the dense result as the structured array the drop-in class
// returns: one packed 9-byte record (score f32, matches i32, overflow u8) per
// pair, row-major, written on the device so that one copy brings the whole
// np.recarray to the host.  Replaces the three device-to-host copies and the
// host-side interleave of np.rec.fromarrays at CudaCosineGreedy.py:201-205.
#include "cs_internal.h"

namespace cs {
namespace {

constexpr int kRecThreads = 256;
constexpr int kRecPerThread = 4;  // 36 bytes: nine aligned 32-bit stores
constexpr int64_t kRecGrid = 65536;

__device__ __forceinline__ void load_rec(const float *score, const int32_t *matches,
                                         const uint8_t *overflow, int64_t cols, int64_t ld,
                                         int64_t k, uint32_t &s, uint32_t &m, uint32_t &o) {
    const int64_t r = k / cols, c = k - r * cols;
    const int64_t x = r * ld + c;
    s = __float_as_uint(score[x]);
    m = (uint32_t)matches[x];
    o = overflow[x];
}

__global__ void __launch_bounds__(kRecThreads)
k_pack_records(const float *score, const int32_t *matches, const uint8_t *overflow, int64_t rows,
               int64_t cols, int64_t ld, uint8_t *records) {
    const int64_t n = rows * cols;
    const int64_t groups = n / kRecPerThread;
    const int64_t stride = (int64_t)gridDim.x * kRecThreads;
    for (int64_t t = blockIdx.x * (int64_t)kRecThreads + threadIdx.x; t < groups; t += stride) {
        uint32_t s[kRecPerThread], m[kRecPerThread], o[kRecPerThread];
#pragma unroll
        for (int u = 0; u < kRecPerThread; ++u)
            load_rec(score, matches, overflow, cols, ld, t * kRecPerThread + u, s[u], m[u], o[u]);
        // record u occupies bytes [9u, 9u + 9): score, matches, overflow
        uint32_t w[9];
        w[0] = s[0];
        w[1] = m[0];
        w[2] = o[0] | (s[1] << 8);
        w[3] = (s[1] >> 24) | (m[1] << 8);
        w[4] = (m[1] >> 24) | (o[1] << 8) | (s[2] << 16);
        w[5] = (s[2] >> 16) | (m[2] << 16);
        w[6] = (m[2] >> 16) | (o[2] << 16) | (s[3] << 24);
        w[7] = (s[3] >> 8) | (m[3] << 24);
        w[8] = (m[3] >> 8) | (o[3] << 24);
        uint32_t *dst = reinterpret_cast<uint32_t *>(records + t * (9 * kRecPerThread));
#pragma unroll
        for (int x = 0; x < 9; ++x) dst[x] = w[x];
    }
    // the last n % 4 records, byte by byte
    const int64_t tail = groups * kRecPerThread;
    if (blockIdx.x == 0 && threadIdx.x < n - tail) {
        const int64_t k = tail + threadIdx.x;
        uint32_t s, m, o;
        load_rec(score, matches, overflow, cols, ld, k, s, m, o);
        uint8_t *d = records + k * 9;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            d[b] = (uint8_t)(s >> (8 * b));
            d[4 + b] = (uint8_t)(m >> (8 * b));
        }
        d[8] = (uint8_t)o;
    }
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_pack_records(const float *score, const int32_t *matches,
                                     const uint8_t *overflow, int64_t rows, int64_t cols,
                                     int64_t ld, uint8_t *records, void *stream) {
    if (rows < 0 || cols < 0 || ld < cols) return fail(CS_ERR_INVALID, "bad record shape");
    if (rows == 0 || cols == 0) return CS_OK;
    if (!score || !matches || !overflow || !records)
        return fail(CS_ERR_INVALID, "record pointer is NULL");
    if (reinterpret_cast<uintptr_t>(records) % 4 != 0)
        return fail(CS_ERR_INVALID, "records must be 4-byte aligned");
    const int64_t groups = rows * cols / kRecPerThread;
    int64_t blocks = (groups + kRecThreads - 1) / kRecThreads;
    blocks = blocks < 1 ? 1 : (blocks > kRecGrid ? kRecGrid : blocks);
    hipLaunchKernelGGL(k_pack_records, dim3((unsigned)blocks), dim3(kRecThreads), 0,
                       static_cast<hipStream_t>(stream), score, matches, overflow, rows, cols, ld,
                       records);
    CS_LAUNCH_CHECK("k_pack_records");
    return CS_OK;
}
