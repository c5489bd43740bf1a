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
constexpr int kTopkChunk = 4096;   // columns of one batch of loads (16 per thread)
constexpr int kTopkLoads = kTopkChunk / kTopkThreads;
constexpr int kTopkBlocks = 512;   // first-pass blocks aimed at: each ranks whole batches
constexpr int kTopkMaxBatches = 16;
constexpr int kTopkStage = kTopkLoads * 64;  // candidates a wave stages: one batch at most
constexpr int kMaxK = 16;

// Ranking of one hit: higher score first, then lower column.  NaN ranks last;
// empty slots are (NaN, INT_MAX) and rank after every real hit.
__device__ __forceinline__ bool hit_before(float sa, int ca, float sb, int cb) {
    const bool na = sa != sa, nb = sb != sb;
    if (na) return nb && ca < cb;
    if (nb) return true;
    return sa > sb || (sa == sb && ca < cb);
}

// The K best hits one thread has seen, best first, in registers.
template <int K>
struct TopList {
    float s[K];
    int c[K];
    __device__ void clear() {
#pragma unroll
        for (int x = 0; x < K; ++x) {
            s[x] = __builtin_nanf("");
            c[x] = INT_MAX;
        }
    }
    __device__ void insert(float v, int cc) {
        if (!hit_before(v, cc, s[K - 1], c[K - 1])) return;
#pragma unroll
        for (int x = 0; x < K; ++x) {  // bubble the new hit into place
            const bool take = hit_before(v, cc, s[x], c[x]);
            const float tv = take ? s[x] : v;
            const int tc = take ? c[x] : cc;
            s[x] = take ? v : s[x];
            c[x] = take ? cc : c[x];
            v = tv;
            cc = tc;
        }
    }
    // drop the head (the caller's list held the wave's best hit)
    __device__ void pop() {
#pragma unroll
        for (int x = 0; x + 1 < K; ++x) {
            s[x] = s[x + 1];
            c[x] = c[x + 1];
        }
        s[K - 1] = __builtin_nanf("");
        c[K - 1] = INT_MAX;
    }
};

template <int J>
__device__ __forceinline__ void best_step(float &bs, int &bc) {
    const float os = xor_lane_float<J>(bs);
    const int oc = xor_lane<J>(bc);
    const bool t = hit_before(os, oc, bs, bc);
    bs = t ? os : bs;
    bc = t ? oc : bc;
}

// Best head over the wave (every lane gets it): a butterfly whose partner
// moves are DPP and swizzle steps, one bpermute (the whole wave must be active).
__device__ __forceinline__ void wave_best(float &bs, int &bc) {
    best_step<1>(bs, bc);
    best_step<2>(bs, bc);
    best_step<4>(bs, bc);
    best_step<8>(bs, bc);
    best_step<16>(bs, bc);
    best_step<32>(bs, bc);
}

// One stage of a descending bitonic sort of one float per lane (no NaN).
template <int SIZE, int J>
__device__ __forceinline__ void fsort_stage(float &x, int lane) {
    const float o = xor_lane_float<J>(x);
    const bool keep_max = ((lane & J) == 0) == ((lane & SIZE) == 0);
    x = keep_max ? fmaxf(x, o) : fminf(x, o);
    if constexpr (J > 1) fsort_stage<SIZE, J / 2>(x, lane);
}
template <int SIZE>
__device__ __forceinline__ void fsort_sizes(float &x, int lane) {
    fsort_stage<SIZE, SIZE / 2>(x, lane);
    if constexpr (SIZE < 64) fsort_sizes<SIZE * 2>(x, lane);
}

// The k-th largest of the wave's 64 values (k uniform, 1 <= k <= 64; no
// NaN): a descending bitonic sort across the lanes, then lane k - 1.  The
// whole wave must be active.
__device__ __forceinline__ float wave_kth_largest(float x, int k) {
    const int lane = threadIdx.x & 63;
    fsort_sizes<2>(x, lane);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k - 1));
}

// Merges the threads' lists of one block; the block's k best (k <= K) end
// in fs[0..k) / fc[0..k) (LDS).  Columns are distinct, so exactly one lane
// owns each real head; empty heads are all alike and pop together.
template <int K>
__device__ void block_merge(TopList<K> &L, int k, float *ws, int *wc, float *fs, int *fc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int x = 0; x < k; ++x) {  // each wave's k best
        float bs = L.s[0];
        int bc = L.c[0];
        wave_best(bs, bc);
        if (lane == 0) {
            ws[wave * K + x] = bs;
            wc[wave * K + x] = bc;
        }
        if (L.c[0] == bc) L.pop();
    }
    __syncthreads();
    if (wave == 0) {  // merge the waves' lists, one candidate per lane
        constexpr int kWaves = kTopkThreads / 64;
        static_assert(kWaves * K <= 64, "one candidate per lane");
        const bool mine = lane < kWaves * K && (lane % K) < k;
        float hs = mine ? ws[lane] : __builtin_nanf("");
        int hc = mine ? wc[lane] : INT_MAX;
        for (int x = 0; x < k; ++x) {
            float bs = hs;
            int bc = hc;
            wave_best(bs, bc);
            if (lane == 0) {
                fs[x] = bs;
                fc[x] = bc;
            }
            if (hc == bc) {
                hs = __builtin_nanf("");
                hc = INT_MAX;
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void write_hits(const float *fs, const int *fc, int k, int64_t row,
                                           int64_t cols, int64_t ld, const int32_t *matches,
                                           float *top_score, int32_t *top_index,
                                           int32_t *top_matches) {
    for (int x = threadIdx.x; x < k; x += blockDim.x) {
        const int c = fc[x];
        const bool real = c != INT_MAX && (int64_t)c < cols;
        top_score[row * k + x] = real ? fs[x] : __builtin_nanf("");
        top_index[row * k + x] = real ? c : -1;
        if (top_matches) top_matches[row * k + x] = real ? matches[row * ld + c] : 0;
    }
}

// Pass 1: block b ranks columns [chunk * ccols, (chunk + 1) * ccols) of one
// row, in batches of kTopkChunk columns with all of a thread's loads of a
// batch issued before its first comparison.  A chunk spans as many batches as
// keep about kTopkBlocks blocks in the grid, so the block merge (the
// latency-bound part) is paid once per chunk, not once per 4096 columns.  One
// chunk per row: the result is final; otherwise the chunk's k best go to the
// parts.
template <int K>
__global__ void __launch_bounds__(kTopkThreads)
k_topk_chunks(const float *__restrict__ score, const int32_t *__restrict__ matches, int64_t rows,
              int64_t cols, int64_t ld, int k, int64_t nchunk, int64_t ccols,
              float *__restrict__ part_s, int *__restrict__ part_c, float *__restrict__ top_score,
              int32_t *__restrict__ top_index, int32_t *__restrict__ top_matches) {
    __shared__ float ws[kTopkThreads / 64 * K], fs[K];
    __shared__ int wc[kTopkThreads / 64 * K], fc[K];
    __shared__ float st_v[kTopkThreads / 64][kTopkStage];
    __shared__ int st_c[kTopkThreads / 64][kTopkStage];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t b = blockIdx.x; b < rows * nchunk; b += gridDim.x) {
        const int64_t row = b / nchunk, chunk = b - row * nchunk;
        const int64_t c0 = chunk * ccols;
        const int64_t c1 = c0 + ccols < cols ? c0 + ccols : cols;
        const float *s = score + row * ld;
        // pass 1: the lane's largest value (NaN skipped), a batch of loads in
        // flight at a time
        float mx = -__builtin_huge_valf();
        for (int64_t base = c0; base < c1; base += kTopkChunk) {
            float v[kTopkLoads];
#pragma unroll
            for (int u = 0; u < kTopkLoads; ++u) {
                const int64_t c = base + u * kTopkThreads + tid;
                v[u] = c < c1 ? s[c] : -__builtin_huge_valf();
            }
#pragma unroll
            for (int u = 0; u < kTopkLoads; ++u) mx = fmaxf(mx, v[u]);
        }
        // the k-th largest lane maximum bounds the block's k-th best from
        // below (it is the k-th of k real values of the block), so only
        // values at or above it can rank; below k real values (-inf) every
        // value is a candidate, NaN included
        const float tau = wave_kth_largest(mx, k);
        const bool all = tau == -__builtin_huge_valf();
        // pass 2: the candidates (the chunk is read again, from the cache)
        // are staged in the wave's LDS list and inserted into the threads'
        // lists from there, one insertion in the code and lanes busy with
        // candidates instead of each lane with its own (few values pass)
        TopList<K> L;
        L.clear();
        int staged = 0;  // uniform
        auto drain = [&]() {
            for (int e = lane; e < staged; e += 64) L.insert(st_v[wave][e], st_c[wave][e]);
            staged = 0;
        };
        for (int64_t base = c0; base < c1; base += kTopkChunk) {
            float v[kTopkLoads];
#pragma unroll
            for (int u = 0; u < kTopkLoads; ++u) {
                const int64_t c = base + u * kTopkThreads + tid;
                v[u] = c < c1 ? s[c] : -__builtin_huge_valf();
            }
#pragma unroll
            for (int u = 0; u < kTopkLoads; ++u) {
                const int64_t c = base + u * kTopkThreads + tid;
                const bool pass = c < c1 && (v[u] >= tau || all);
                const unsigned long long m = __ballot(pass);
                if (pass) {
                    const int at = staged + __popcll(m & ((1ull << lane) - 1ull));
                    st_v[wave][at] = v[u];
                    st_c[wave][at] = (int)c;
                }
                staged += __popcll(m);
            }
            if (staged > kTopkStage - kTopkLoads * 64) drain();  // uniform: room for a batch
        }
        drain();
        block_merge<K>(L, k, ws, wc, fs, fc);
        if (nchunk == 1) {
            write_hits(fs, fc, k, row, cols, ld, matches, top_score, top_index, top_matches);
        } else {
            for (int x = tid; x < k; x += kTopkThreads) {
                part_s[b * K + x] = fs[x];
                part_c[b * K + x] = fc[x];
            }
        }
        __syncthreads();
    }
}

// Pass 2: one block per row merges the row's nchunk x k partial hits.
template <int K>
__global__ void __launch_bounds__(kTopkThreads)
k_topk_merge(const float *__restrict__ part_s, const int *__restrict__ part_c,
             const int32_t *__restrict__ matches, int64_t rows, int64_t cols, int64_t ld, int k,
             int64_t nchunk, float *__restrict__ top_score, int32_t *__restrict__ top_index,
             int32_t *__restrict__ top_matches) {
    __shared__ float ws[kTopkThreads / 64 * K], fs[K];
    __shared__ int wc[kTopkThreads / 64 * K], fc[K];
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        TopList<K> L;
        L.clear();
        const int64_t n = nchunk * K;
        for (int64_t e = threadIdx.x; e < n; e += kTopkThreads) {
            if ((e % K) < k) L.insert(part_s[row * n + e], part_c[row * n + e]);
        }
        block_merge<K>(L, k, ws, wc, fs, fc);
        write_hits(fs, fc, k, row, cols, ld, matches, top_score, top_index, top_matches);
        __syncthreads();
    }
}

int topk_width(int k) { return k <= 4 ? 4 : (k <= 10 ? 10 : 16); }

// Batches of kTopkChunk columns per row (the most first-pass chunks a row can
// have: the workspace holds that many partial lists per row).
int64_t topk_chunks(int64_t cols) {
    const int64_t n = (cols + kTopkChunk - 1) / kTopkChunk;
    return n < 1 ? 1 : n;
}

// Columns of one first-pass chunk for a rows x cols block: whole batches,
// about kTopkBlocks chunks in all, at most kTopkMaxBatches batches each.
int64_t topk_chunk_cols(int64_t rows, int64_t cols) {
    int64_t m = rows * topk_chunks(cols) / kTopkBlocks;
    m = m < 1 ? 1 : (m > kTopkMaxBatches ? kTopkMaxBatches : m);
    return m * kTopkChunk;
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

extern "C" cs_status cs_topk_rows_workspace_size(int64_t rows, int64_t cols, int k,
                                                 size_t *bytes) {
    if (rows < 0 || cols < 0) return fail(CS_ERR_INVALID, "bad top-k shape");
    if (k < 1 || k > kMaxK) return fail(CS_ERR_INVALID, "k must be in [1, %d]", kMaxK);
    if (!bytes) return fail(CS_ERR_INVALID, "bytes is NULL");
    const int64_t nchunk = topk_chunks(cols);
    const int64_t n = nchunk > 1 ? rows * nchunk * topk_width(k) : 0;
    *bytes = align_up((size_t)n * sizeof(float), 256) + align_up((size_t)n * sizeof(int), 256);
    return CS_OK;
}

template <int K>
void launch_topk(const float *score, const int32_t *matches, int64_t rows, int64_t cols,
                 int64_t ld, int k, float *top_score, int32_t *top_index, int32_t *top_matches,
                 float *part_s, int *part_c, int64_t nchunk, int64_t ccols, hipStream_t hs) {
    hipLaunchKernelGGL(k_topk_chunks<K>, dim3(grid_for(rows * nchunk, 1 << 20)),
                       dim3(kTopkThreads), 0, hs, score, matches, rows, cols, ld, k, nchunk,
                       ccols, part_s, part_c, top_score, top_index, top_matches);
    if (nchunk > 1)
        hipLaunchKernelGGL(k_topk_merge<K>, dim3(grid_for(rows, 65536)), dim3(kTopkThreads), 0,
                           hs, part_s, part_c, matches, rows, cols, ld, k, nchunk, top_score,
                           top_index, top_matches);
}

extern "C" cs_status cs_topk_rows(const float *score, const int32_t *matches, int64_t rows,
                                  int64_t cols, int64_t ld, int k, float *top_score,
                                  int32_t *top_index, int32_t *top_matches, void *workspace,
                                  size_t workspace_bytes, void *stream) {
    if (rows < 0 || cols < 0 || ld < cols) return fail(CS_ERR_INVALID, "bad top-k shape");
    if (k < 1 || k > kMaxK) return fail(CS_ERR_INVALID, "k must be in [1, %d]", kMaxK);
    if (cols > INT_MAX - 1) return fail(CS_ERR_UNSUPPORTED, "more than 2^31 columns");
    if (rows == 0) return CS_OK;
    if (!score || !top_score || !top_index || (top_matches && !matches))
        return fail(CS_ERR_INVALID, "top-k pointer is NULL");
    size_t need = 0;
    cs_status st = cs_topk_rows_workspace_size(rows, cols, k, &need);
    if (st != CS_OK) return st;
    if (need > 0 && (!workspace || workspace_bytes < need))
        return fail(CS_ERR_WORKSPACE, "workspace of %zu bytes < %zu required", workspace_bytes,
                    need);
    // the workspace holds topk_chunks(cols) partial lists per row; a launch
    // uses nchunk <= that many
    const int64_t most = topk_chunks(cols);
    const int64_t n = most > 1 ? rows * most * topk_width(k) : 0;
    Carver c(workspace);
    float *part_s = c.take<float>(n);
    int *part_c = c.take<int>(n);
    const int64_t ccols = topk_chunk_cols(rows, cols);
    const int64_t nchunk = (cols + ccols - 1) / ccols < 1 ? 1 : (cols + ccols - 1) / ccols;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    switch (topk_width(k)) {
        case 4:
            launch_topk<4>(score, matches, rows, cols, ld, k, top_score, top_index, top_matches,
                           part_s, part_c, nchunk, ccols, hs);
            break;
        case 10:
            launch_topk<10>(score, matches, rows, cols, ld, k, top_score, top_index, top_matches,
                            part_s, part_c, nchunk, ccols, hs);
            break;
        default:
            launch_topk<16>(score, matches, rows, cols, ld, k, top_score, top_index, top_matches,
                            part_s, part_c, nchunk, ccols, hs);
            break;
    }
    CS_LAUNCH_CHECK("k_topk");
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
