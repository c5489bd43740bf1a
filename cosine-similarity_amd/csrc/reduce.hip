// reduce.hip -- reductions over a dense score block: per-row top-k and
// threshold compaction into COO.
//
// Top-k replaces the host-side "best hits of a query" the reference's users
// take with matchms Scores.scores_by_query(sort=True)
// (tests/usecase/test_tutorial_pesticide.py:38-41, :64-67).  Compaction
// replaces the per-tile mask / nonzero / D2H of CudaCosineGreedy.py:186-200.
#include "cs_internal.h"

namespace cs {
namespace {

constexpr int kTopkThreads = 256;
constexpr int kMaxK = 16;

// Ranking of one hit: higher score first, then lower column.  NaN ranks last.
__device__ __forceinline__ bool hit_before(float sa, int ca, float sb, int cb) {
    const bool na = sa != sa, nb = sb != sb;
    if (na) return nb && ca < cb;
    if (nb) return true;
    return sa > sb || (sa == sb && ca < cb);
}

template <int K>
__global__ void __launch_bounds__(kTopkThreads)
k_topk_rows(const float *__restrict__ score, const int32_t *__restrict__ matches, int64_t rows,
            int64_t cols, int64_t ld, int k, float *__restrict__ top_score,
            int32_t *__restrict__ top_index, int32_t *__restrict__ top_matches) {
    __shared__ float ls[kTopkThreads * K];
    __shared__ int lc[kTopkThreads * K];
    const int tid = threadIdx.x;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const float *s = score + row * ld;
        float bs[K];
        int bc[K];
#pragma unroll
        for (int x = 0; x < K; ++x) {
            bs[x] = __builtin_nanf("");
            bc[x] = INT_MAX;
        }
        for (int64_t c = tid; c < cols; c += kTopkThreads) {
            float v = s[c];
            int cc = (int)c;
            if (!hit_before(v, cc, bs[K - 1], bc[K - 1])) continue;
#pragma unroll
            for (int x = 0; x < K; ++x) {  // bubble the new hit into place
                const bool take = hit_before(v, cc, bs[x], bc[x]);
                const float tv = take ? bs[x] : v;
                const int tc = take ? bc[x] : cc;
                bs[x] = take ? v : bs[x];
                bc[x] = take ? cc : bc[x];
                v = tv;
                cc = tc;
            }
        }
#pragma unroll
        for (int x = 0; x < K; ++x) {
            ls[x * kTopkThreads + tid] = bs[x];
            lc[x * kTopkThreads + tid] = bc[x];
        }
        __syncthreads();
        // bitonic sort of the kTopkThreads * K candidates by hit_before
        const int n = kTopkThreads * K;
        for (int size = 2; size <= n; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < n / 2; t += kTopkThreads) {
                    const int lo = 2 * t - (t & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const float a = ls[lo], b = ls[hi];
                    const int ca = lc[lo], cb = lc[hi];
                    const bool sw = up ? hit_before(b, cb, a, ca) : hit_before(a, ca, b, cb);
                    if (sw) {
                        ls[lo] = b;
                        ls[hi] = a;
                        lc[lo] = cb;
                        lc[hi] = ca;
                    }
                }
                __syncthreads();
            }
        }
        for (int x = tid; x < k; x += kTopkThreads) {
            const int c = lc[x];
            const bool real = c != INT_MAX && (int64_t)c < cols;
            top_score[row * k + x] = real ? ls[x] : __builtin_nanf("");
            top_index[row * k + x] = real ? c : -1;
            if (top_matches) top_matches[row * k + x] = real ? matches[row * ld + c] : 0;
        }
        __syncthreads();
    }
}

// Per-row count of score >= threshold (NaN never passes).
__global__ void __launch_bounds__(256)
k_count_rows(const float *__restrict__ score, int64_t rows, int64_t cols, int64_t ld,
             float threshold, int32_t *__restrict__ counts) {
    __shared__ int part[256];
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        int c = 0;
        for (int64_t q = threadIdx.x; q < cols; q += 256) c += score[row * ld + q] >= threshold;
        part[threadIdx.x] = c;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) counts[row] = part[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) k_scan_rows(const int32_t *counts, int64_t n, int64_t *out) {
    __shared__ int64_t part[1024];
    const int tid = threadIdx.x;
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t b = tid * chunk;
    const int64_t e = b + chunk < n ? b + chunk : n;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t t = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += t;
        __syncthreads();
    }
    int64_t run = tid > 0 ? part[tid - 1] : 0;
    for (int64_t i = b; i < e; ++i) {
        out[i] = run;
        run += counts[i];
    }
    if (tid == 1023) out[n] = part[1023];
}

// Writes the passing entries of each row in column order at row_offsets[row].
__global__ void __launch_bounds__(64)
k_compact_rows(const float *__restrict__ score, const int32_t *__restrict__ matches,
               const uint8_t *__restrict__ overflow, int64_t rows, int64_t cols, int64_t ld,
               float threshold, int64_t row0, const int64_t *__restrict__ row_offsets,
               int64_t capacity, int64_t *__restrict__ out_row, int64_t *__restrict__ out_col,
               float *__restrict__ out_score, int32_t *__restrict__ out_matches,
               uint8_t *__restrict__ out_overflow) {
    const int lane = threadIdx.x;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        int64_t pos = row_offsets[row];
        for (int64_t base = 0; base < cols; base += 64) {
            const int64_t q = base + lane;
            const int64_t o = row * ld + q;
            const bool keep = q < cols && score[o] >= threshold;
            const unsigned long long mask = __ballot(keep);
            if (keep) {
                const int64_t at = pos + __popcll(mask & ((1ull << lane) - 1ull));
                if (at < capacity) {
                    out_row[at] = row0 + row;
                    out_col[at] = q;
                    out_score[at] = score[o];
                    if (out_matches) out_matches[at] = matches[o];
                    if (out_overflow) out_overflow[at] = overflow[o];
                }
            }
            pos += __popcll(mask);
        }
    }
}

unsigned grid_for(int64_t n, int64_t cap) { return (unsigned)(n < 1 ? 1 : (n > cap ? cap : n)); }

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_topk_rows(const float *score, const int32_t *matches, int64_t rows,
                                  int64_t cols, int64_t ld, int k, float *top_score,
                                  int32_t *top_index, int32_t *top_matches, void *stream) {
    if (rows < 0 || cols < 0 || ld < cols) return fail(CS_ERR_INVALID, "bad top-k shape");
    if (k < 1 || k > kMaxK) return fail(CS_ERR_INVALID, "k must be in [1, %d]", kMaxK);
    if (cols > INT_MAX - 1) return fail(CS_ERR_UNSUPPORTED, "more than 2^31 columns");
    if (rows == 0) return CS_OK;
    if (!score || !top_score || !top_index || (top_matches && !matches))
        return fail(CS_ERR_INVALID, "top-k pointer is NULL");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const unsigned g = grid_for(rows, 65536);
    if (k <= 4)
        hipLaunchKernelGGL(k_topk_rows<4>, dim3(g), dim3(kTopkThreads), 0, hs, score, matches, rows,
                           cols, ld, k, top_score, top_index, top_matches);
    else if (k <= 8)
        hipLaunchKernelGGL(k_topk_rows<8>, dim3(g), dim3(kTopkThreads), 0, hs, score, matches, rows,
                           cols, ld, k, top_score, top_index, top_matches);
    else
        hipLaunchKernelGGL(k_topk_rows<16>, dim3(g), dim3(kTopkThreads), 0, hs, score, matches,
                           rows, cols, ld, k, top_score, top_index, top_matches);
    CS_LAUNCH_CHECK("k_topk_rows");
    return CS_OK;
}

extern "C" cs_status cs_threshold_compact(const float *score, const int32_t *matches,
                                          const uint8_t *overflow, int64_t rows, int64_t cols,
                                          int64_t ld, float threshold, int64_t row0,
                                          int64_t *row_offsets, int32_t *row_counts,
                                          int64_t capacity, int64_t *out_row, int64_t *out_col,
                                          float *out_score, int32_t *out_matches,
                                          uint8_t *out_overflow, void *stream) {
    if (rows < 0 || cols < 0 || ld < cols) return fail(CS_ERR_INVALID, "bad compaction shape");
    if (rows == 0) return CS_OK;
    if (!score || !row_offsets || !row_counts)
        return fail(CS_ERR_INVALID, "score/row_offsets/row_counts must not be NULL");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const unsigned g = grid_for(rows, 65536);
    if (!out_row) {  // count phase
        hipLaunchKernelGGL(k_count_rows, dim3(g), dim3(256), 0, hs, score, rows, cols, ld,
                           threshold, row_counts);
        CS_LAUNCH_CHECK("k_count_rows");
        hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, hs, row_counts, rows, row_offsets);
        CS_LAUNCH_CHECK("k_scan_rows");
        return CS_OK;
    }
    if (!out_col || !out_score || (out_matches && !matches) || (out_overflow && !overflow))
        return fail(CS_ERR_INVALID, "compaction output pointer is NULL");
    hipLaunchKernelGGL(k_compact_rows, dim3(g), dim3(64), 0, hs, score, matches, overflow, rows,
                       cols, ld, threshold, row0, row_offsets, capacity, out_row, out_col,
                       out_score, out_matches, out_overflow);
    CS_LAUNCH_CHECK("k_compact_rows");
    return CS_OK;
}
