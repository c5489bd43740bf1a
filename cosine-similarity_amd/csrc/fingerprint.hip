// fingerprint.hip -- Jaccard / Dice / cosine between binary fingerprints on
// gfx950 int8 matrix cores.
//
// Replaces the torch float32 GEMM + elementwise epilogues of
// cudams/similarity/vector_similarity_functions.py:12-47 (jaccard),
// :50-91 (dice) and :94-119 (cosine).
//
// Pipeline (asynchronous on the caller's stream):
//   1. k_fp_expand   bit-packed uint32 rows -> int8 0/1 rows, zero padded to a
//                    multiple of 128 rows and 64 bytes of K
//   2. k_fp_popcount |row| per fingerprint (exact int32)
//   3. k_fp_gemm     128 x 128 output tile per 256-thread workgroup, 2 x 2
//                    waves of 64 x 64, v_mfma_i32_16x16x64_i8 (exact int32
//                    intersections), K staged through LDS in 64-byte slices,
//                    measure + missing-fingerprint fill fused in the epilogue.
#include <limits.h>

#include "cs_internal.h"

namespace cs {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kTile = 128;
constexpr int kKc = 64;          // bytes of K per MFMA step
constexpr int kThreads = 256;
constexpr int kRowStride = 80;   // LDS bytes per row: 64 + 16 pad keeps ds_read_b128 conflict-free

__global__ void k_fp_expand(const uint32_t *__restrict__ words, int64_t rows, int64_t rows_pad,
                            int64_t n_words, int64_t k_bytes, uint32_t *__restrict__ out) {
    // out row x holds k_bytes bytes = k_bytes / 4 dwords; each thread writes one dword
    const int64_t dw_per_row = k_bytes / 4;
    const int64_t total = rows_pad * dw_per_row;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = e / dw_per_row;
        const int64_t d = e - x * dw_per_row;  // bytes 4d .. 4d+3 = bits 4d .. 4d+3
        const int64_t w = d >> 3;              // 8 dwords per source word
        uint32_t v = 0u;
        if (x < rows && w < n_words) {
            const uint32_t nib = (words[x * n_words + w] >> ((d & 7) * 4)) & 0xfu;
            v = (nib * 0x00204081u) & 0x01010101u;  // bit b of the nibble -> byte b
        }
        out[e] = v;
    }
}

__global__ void k_fp_popcount(const uint32_t *__restrict__ words, int64_t rows, int64_t n_words,
                              int32_t *__restrict__ pop) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < rows;
         x += (int64_t)gridDim.x * blockDim.x) {
        int32_t c = 0;
        for (int64_t w = 0; w < n_words; ++w) c += __popc(words[x * n_words + w]);
        pop[x] = c;
    }
}

__device__ __forceinline__ float measure_value(int measure, int32_t inter, int32_t a, int32_t b) {
    if (measure == CS_JACCARD) {
        const int32_t uni = a + b - inter;
        return uni ? (float)((double)inter / (double)uni) : 0.0f;
    }
    if (measure == CS_DICE) {
        const int32_t den = a + b;
        return den ? (float)(2.0 * ((double)inter / (double)den)) : 0.0f;
    }
    return (a && b) ? (float)((double)inter / sqrt((double)a * (double)b)) : 0.0f;
}

__global__ void __launch_bounds__(kThreads)
k_fp_gemm(const int8_t *__restrict__ A, const int8_t *__restrict__ B, int64_t k_bytes,
          const int32_t *__restrict__ pa, const int32_t *__restrict__ pb, int64_t nR, int64_t nQ,
          const uint8_t *__restrict__ ref_valid, const uint8_t *__restrict__ query_valid,
          float empty_score, int measure, float *__restrict__ out,
          int32_t *__restrict__ inter_out, int64_t ld) {
    __shared__ __align__(16) int8_t lA[kTile * kRowStride];
    __shared__ __align__(16) int8_t lB[kTile * kRowStride];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t r0 = (int64_t)blockIdx.y * kTile;
    const int64_t q0 = (int64_t)blockIdx.x * kTile;

    v4i acc[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = v4i{0, 0, 0, 0};

    // Global -> LDS mapping: 128 rows x 64 bytes = 512 chunks of 16 bytes per
    // matrix; thread tid moves chunks tid and tid + 256 of A and of B.
    const int c0 = tid, c1 = tid + kThreads;
    const int row0 = c0 >> 2, part0 = c0 & 3, row1 = c1 >> 2, part1 = c1 & 3;
    const int8_t *a0 = A + (r0 + row0) * k_bytes + 16 * part0;
    const int8_t *a1 = A + (r0 + row1) * k_bytes + 16 * part1;
    const int8_t *b0 = B + (q0 + row0) * k_bytes + 16 * part0;
    const int8_t *b1 = B + (q0 + row1) * k_bytes + 16 * part1;

    v4i ga0 = *reinterpret_cast<const v4i *>(a0);
    v4i ga1 = *reinterpret_cast<const v4i *>(a1);
    v4i gb0 = *reinterpret_cast<const v4i *>(b0);
    v4i gb1 = *reinterpret_cast<const v4i *>(b1);

    const int frag_row = lane & 15, frag_k = 16 * (lane >> 4);
    for (int64_t kc = 0; kc < k_bytes; kc += kKc) {
        *reinterpret_cast<v4i *>(&lA[row0 * kRowStride + 16 * part0]) = ga0;
        *reinterpret_cast<v4i *>(&lA[row1 * kRowStride + 16 * part1]) = ga1;
        *reinterpret_cast<v4i *>(&lB[row0 * kRowStride + 16 * part0]) = gb0;
        *reinterpret_cast<v4i *>(&lB[row1 * kRowStride + 16 * part1]) = gb1;
        __syncthreads();
        if (kc + kKc < k_bytes) {  // prefetch the next slice while the MFMAs run
            ga0 = *reinterpret_cast<const v4i *>(a0 + kc + kKc);
            ga1 = *reinterpret_cast<const v4i *>(a1 + kc + kKc);
            gb0 = *reinterpret_cast<const v4i *>(b0 + kc + kKc);
            gb1 = *reinterpret_cast<const v4i *>(b1 + kc + kKc);
        }
        v4i fa[4], fb[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
            fa[bi] = *reinterpret_cast<const v4i *>(
                &lA[(wr * 64 + bi * 16 + frag_row) * kRowStride + frag_k]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            fb[bj] = *reinterpret_cast<const v4i *>(
                &lB[(wc * 64 + bj * 16 + frag_row) * kRowStride + frag_k]);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] =
                    __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[bi], fb[bj], acc[bi][bj], 0, 0, 0);
        __syncthreads();
    }

    // Epilogue.  16x16 accumulator layout: lane l, element e holds
    // row 4 * (l / 16) + e, column l % 16 of the block.
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t r = r0 + wr * 64 + bi * 16 + 4 * (lane >> 4) + e;
            if (r >= nR) continue;
            const int32_t a = pa[r];
            const bool rv = !ref_valid || ref_valid[r];
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                const int64_t q = q0 + wc * 64 + bj * 16 + (lane & 15);
                if (q >= nQ) continue;
                const int32_t inter = acc[bi][bj][e];
                const bool ok = rv && (!query_valid || query_valid[q]);
                out[r * ld + q] = ok ? measure_value(measure, inter, a, pb[q]) : empty_score;
                if (inter_out) inter_out[r * ld + q] = inter;
            }
        }
    }
}

struct FpLayout {
    int8_t *A, *B;
    int32_t *pa, *pb;
    int64_t rows_a, rows_b, k_bytes;
    size_t bytes;
};

FpLayout fp_layout(void *base, int64_t nR, int64_t nQ, int64_t n_words) {
    FpLayout L{};
    Carver c(base);
    L.rows_a = (nR + kTile - 1) / kTile * kTile;
    L.rows_b = (nQ + kTile - 1) / kTile * kTile;
    L.k_bytes = (n_words * 32 + kKc - 1) / kKc * kKc;
    L.A = c.take<int8_t>(L.rows_a * L.k_bytes);
    L.B = c.take<int8_t>(L.rows_b * L.k_bytes);
    L.pa = c.take<int32_t>(nR > 0 ? nR : 1);
    L.pb = c.take<int32_t>(nQ > 0 ? nQ : 1);
    L.bytes = align_up(c.used, 256);
    return L;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_fingerprint_workspace_size(int64_t n_refs, int64_t n_queries,
                                                   int64_t n_words, size_t *bytes) {
    if (n_refs < 0 || n_queries < 0 || n_words < 0 || !bytes)
        return fail(CS_ERR_INVALID, "invalid fingerprint shape");
    *bytes = fp_layout(nullptr, n_refs, n_queries, n_words).bytes;
    return CS_OK;
}

extern "C" cs_status cs_fingerprint_similarity(int measure, const uint32_t *ref_words,
                                               int64_t n_refs, const uint32_t *query_words,
                                               int64_t n_queries, int64_t n_words,
                                               const uint8_t *ref_valid,
                                               const uint8_t *query_valid, float empty_score,
                                               float *out, int32_t *intersections, int64_t ld,
                                               void *workspace, size_t workspace_bytes,
                                               void *stream) {
    if (measure != CS_JACCARD && measure != CS_DICE && measure != CS_COSINE)
        return fail(CS_ERR_INVALID, "unknown measure %d", measure);
    if (n_refs < 0 || n_queries < 0 || n_words < 0)
        return fail(CS_ERR_INVALID, "negative fingerprint shape");
    if (n_refs == 0 || n_queries == 0) return CS_OK;
    if (n_words == 0 || !ref_words || !query_words || !out)
        return fail(CS_ERR_INVALID, "fingerprint pointers must not be NULL and n_words > 0");
    if (ld < n_queries) return fail(CS_ERR_INVALID, "ld < number of queries");
    if (n_words * 32 > (int64_t)INT_MAX / 2)
        return fail(CS_ERR_UNSUPPORTED, "fingerprints longer than 2^30 bits");
    const FpLayout need = fp_layout(nullptr, n_refs, n_queries, n_words);
    if (!workspace || workspace_bytes < need.bytes)
        return fail(CS_ERR_WORKSPACE, "workspace of %zu bytes < %zu required", workspace_bytes,
                    need.bytes);
    const FpLayout L = fp_layout(workspace, n_refs, n_queries, n_words);
    hipStream_t hs = static_cast<hipStream_t>(stream);
    auto blocks = [](int64_t n, int per) {
        int64_t b = (n + per - 1) / per;
        return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
    };
    hipLaunchKernelGGL(k_fp_expand, dim3(blocks(L.rows_a * L.k_bytes / 4, 256)), dim3(256), 0, hs,
                       ref_words, n_refs, L.rows_a, n_words, L.k_bytes,
                       reinterpret_cast<uint32_t *>(L.A));
    CS_LAUNCH_CHECK("k_fp_expand");
    hipLaunchKernelGGL(k_fp_expand, dim3(blocks(L.rows_b * L.k_bytes / 4, 256)), dim3(256), 0, hs,
                       query_words, n_queries, L.rows_b, n_words, L.k_bytes,
                       reinterpret_cast<uint32_t *>(L.B));
    CS_LAUNCH_CHECK("k_fp_expand");
    hipLaunchKernelGGL(k_fp_popcount, dim3(blocks(n_refs, 256)), dim3(256), 0, hs, ref_words,
                       n_refs, n_words, L.pa);
    CS_LAUNCH_CHECK("k_fp_popcount");
    hipLaunchKernelGGL(k_fp_popcount, dim3(blocks(n_queries, 256)), dim3(256), 0, hs,
                       query_words, n_queries, n_words, L.pb);
    CS_LAUNCH_CHECK("k_fp_popcount");
    const int64_t gx = L.rows_b / kTile, gy = L.rows_a / kTile;
    if (gx > INT_MAX || gy > 65535)
        return fail(CS_ERR_UNSUPPORTED, "fingerprint grid too large");
    hipLaunchKernelGGL(k_fp_gemm, dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, hs, L.A,
                       L.B, L.k_bytes, L.pa, L.pb, n_refs, n_queries, ref_valid, query_valid,
                       empty_score, measure, out, intersections, ld);
    CS_LAUNCH_CHECK("k_fp_gemm");
    return CS_OK;
}
