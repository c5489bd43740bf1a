// fingerprint.hip -- Jaccard / Dice / cosine between binary fingerprints on
// gfx950 matrix cores.
//
// Replaces the torch float32 GEMM + elementwise epilogues of
// cudams/similarity/vector_similarity_functions.py:12-47 (jaccard),
// :50-91 (dice) and :94-119 (cosine).
//
// One kernel, no workspace.  The packed uint32 rows are read as they are;
// each 4-bit nibble plane of a word becomes 8 fp4 (e2m1) values with one
// VALU op:
//     plane 0: w & 0x11111111 -> 0.5     plane 1: w & 0x22222222 -> 1.0
//     plane 2: w & 0x44444444 -> 2.0     plane 3: (w >> 1) & 0x44444444 -> 2.0
// (a cleared bit gives 0).  v_mfma_scale_f32_16x16x128_f8f6f4 multiplies
// both operands by a power-of-two block scale (2, 1, 1/2, 1/2 for the four
// planes), so every set bit is exactly 1.0 and the float32 accumulator holds
// the exact intersection count (integers below 2^24).  Reference and query
// operands use the same bit -> element mapping, so the element order the
// matrix core assigns to K does not matter.
//
// Workgroup: 256 threads, 2 x 2 waves, each wave a 64-reference x 64-query
// tile = 4 x 4 blocks of 16 x 16; two workgroups per CU (amdgpu_waves_per_eu
// 2, no LDS).  The grid is persistent (2 workgroups per CU) and walks the
// 128 x 128 block tiles.  Each lane loads, for each of its 8 fragments (4
// reference, 4 query rows), 4 words (128 bits) per 512-bit step straight
// from global memory (the packed rows are small and stay in L2), one step
// ahead; the last step of a tile prefetches the next tile's first step.  Row
// popcounts come from v_bcnt on the same words.  The epilogue writes 4
// consecutive queries per lane (one 16-byte non-temporal store) with the
// measure computed in float32 (jaccard, dice: the correctly rounded quotient
// of exact integers, which equals the float32 rounding of the float64
// quotient; cosine: the count times a reciprocal square root).
//
// On gfx950 VALU instructions and MFMAs of one SIMD do not overlap, so the
// kernel is priced as MFMA cycles + VALU cycles: every VALU op in the K loop
// (plane masks, popcounts) and in the epilogue (the division) counts.
#include <limits.h>

#include <algorithm>
#include <atomic>

#include "cs_internal.h"

namespace cs {
namespace {

typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

#ifndef CS_FP_REF_FRAGS
#define CS_FP_REF_FRAGS 4
#endif
#ifndef CS_FP_QUERY_FRAGS
#define CS_FP_QUERY_FRAGS 4
#endif
constexpr int kRF = CS_FP_REF_FRAGS;     // 16-reference fragments per wave
constexpr int kQF = CS_FP_QUERY_FRAGS;   // 16-query fragments per wave
constexpr int kRefTile = 2 * 16 * kRF;   // references per workgroup (2 x 2 waves)
constexpr int kQueryTile = 2 * 16 * kQF;
constexpr int kThreads = 256;
constexpr int kStepWords = 16;   // 512 bits of K per step (4 lane groups x 4 words)
constexpr int64_t kMaxBits = 1 << 23;  // counts and |a| + |b| stay exact in float32

template <bool kFull>
__device__ __forceinline__ v4u load_words(const uint32_t *__restrict__ row, int64_t w,
                                          int64_t n_words) {
    if (kFull) return *reinterpret_cast<const v4u *>(row + w);
    v4u v;
    v.x = w < n_words ? row[w] : 0u;
    v.y = w + 1 < n_words ? row[w + 1] : 0u;
    v.z = w + 2 < n_words ? row[w + 2] : 0u;
    v.w = w + 3 < n_words ? row[w + 3] : 0u;
    return v;
}

template <bool kFull>
__device__ __forceinline__ void load_step(const uint32_t *const (&rrow)[kRF],
                                          const uint32_t *const (&qrow)[kQF], int64_t w,
                                          int64_t n_words, v4u (&r)[kRF], v4u (&q)[kQF]) {
#pragma unroll
    for (int f = 0; f < kRF; ++f) r[f] = load_words<kFull>(rrow[f], w, n_words);
#pragma unroll
    for (int f = 0; f < kQF; ++f) q[f] = load_words<kFull>(qrow[f], w, n_words);
}

template <int P>
__device__ __forceinline__ v8i plane(v4u x) {
    v4u y;
    if (P == 0) y = x & 0x11111111u;
    else if (P == 1) y = x & 0x22222222u;
    else if (P == 2) y = x & 0x44444444u;
    else y = (x >> 1u) & 0x44444444u;
    return v8i{(int)y.x, (int)y.y, (int)y.z, (int)y.w, 0, 0, 0, 0};
}

// E8M0 block scale that turns plane P's nonzero value into 1.0 (the same
// scale on both operands: (v * s)^2 = 1 for v * s = 1).
template <int P>
constexpr int plane_scale() { return P == 0 ? 128 : (P == 1 ? 127 : 126); }

// acc + popcount(x) as one v_bcnt_u32_b32 per word (the compiler otherwise
// splits the chain into separate counts and adds).
__device__ __forceinline__ uint32_t popc4(v4u x, uint32_t acc) {
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x.x));
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x.y));
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x.z));
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(acc) : "v"(x.w));
    return acc;
}


// Correctly rounded n / d for integers 0 <= n <= 2^24, 1 <= d <= 2^24 held
// exactly in float32.  This is the IEEE float32 division sequence (reciprocal,
// one Newton step on it, two residual corrections of the quotient, the last
// FMA rounding once) without the operand scaling and special-case fixup,
// which are identities on this domain: no operand or intermediate is
// denormal, infinite or NaN, and n = 0 gives +0 exactly.
__device__ __forceinline__ float div_rn(float n, float d) {
    float r = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, r, 1.0f);
    r = fmaf(e, r, r);
    float q = n * r;
    float t = fmaf(-d, q, n);
    q = fmaf(t, r, q);
    t = fmaf(-d, q, n);
    return fmaf(t, r, q);
}

// n / d for the (intersection, denominator) pairs of fingerprints of at most
// kShortDivBits bits: the hardware reciprocal, the quotient, and one residual
// correction rounding once.  On this domain the result is the correctly
// rounded quotient: the GPU tests check every such pair against the float32
// rounding of the float64 quotient (fingerprints with prefix bit patterns give
// intersection min(a, b) with the jaccard denominator max(a, b) and the dice
// denominator a + b, for all a, b <= 4096).
constexpr int64_t kShortDivBits = 4096;  // fingerprints up to this length use div_short
__device__ __forceinline__ float div_short(float n, float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float q = n * r;
    const float t = fmaf(-d, q, n);
    return fmaf(t, r, q);
}

// Measure from the exact counts (float32 integers up to 2^23 bits, so every
// sum below is exact).  The caller passes a and b raised to at least 0.5
// (count_floor): a count is 0 only for an empty fingerprint, and then inter
// is 0 too, so the denominator stays positive and the quotient is +0 exactly,
// the reference's nan_to_num 0, without a per-pair clamp; nonzero counts are
// unchanged.  Cosine, whose bar is a tolerance (the reference computes it in
// float32, vector_similarity_functions.py:116), is the count times the
// hardware reciprocal square root of a * b: within a few float32 ulps of the
// float64 value.
template <int kMeasure, bool kShort>
__device__ __forceinline__ float measure_value(float inter, float a, float b) {
    if (kMeasure == CS_JACCARD) {
        const float u = (a - inter) + b;
        return kShort ? div_short(inter, u) : div_rn(inter, u);
    }
    if (kMeasure == CS_DICE) {
        const float u = a + b;
        return 2.0f * (kShort ? div_short(inter, u) : div_rn(inter, u));
    }
    return inter * __builtin_amdgcn_rsqf(a * b);
}

// A row count for measure_value: the count, or 0.5 for an empty fingerprint.
__device__ __forceinline__ float count_floor(uint32_t c) { return fmaxf((float)c, 0.5f); }

template <int P>
__device__ __forceinline__ void plane_mfma(const v4u (&r)[kRF], const v4u (&q)[kQF],
                                           v4f (&acc)[kQF][kRF]) {
    v8i br[kRF];
#pragma unroll
    for (int f = 0; f < kRF; ++f) br[f] = plane<P>(r[f]);
#pragma unroll
    for (int qf = 0; qf < kQF; ++qf) {
        const v8i aq = plane<P>(q[qf]);
#pragma unroll
        for (int rf = 0; rf < kRF; ++rf)
            acc[qf][rf] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                aq, br[rf], acc[qf][rf], 4, 4, 0, plane_scale<P>(), 0, plane_scale<P>());
    }
}

// One 512-bit step: popcounts, then the 4 planes x kQF x kRF blocks.
__device__ __forceinline__ void fp_step(const v4u (&r)[kRF], const v4u (&q)[kQF],
                                        v4f (&acc)[kQF][kRF], uint32_t (&pr)[kRF],
                                        uint32_t (&pq)[kQF]) {
#pragma unroll
    for (int f = 0; f < kRF; ++f) pr[f] = popc4(r[f], pr[f]);
#pragma unroll
    for (int f = 0; f < kQF; ++f) pq[f] = popc4(q[f], pq[f]);
    // The barriers keep the scheduler from expanding every plane up front
    // (32 extra registers per plane).
    plane_mfma<0>(r, q, acc);
    __builtin_amdgcn_sched_barrier(0);
    plane_mfma<1>(r, q, acc);
    __builtin_amdgcn_sched_barrier(0);
    plane_mfma<2>(r, q, acc);
    __builtin_amdgcn_sched_barrier(0);
    plane_mfma<3>(r, q, acc);
}

// Wave-level row bases of block tile `tile` and the row pointers the lane
// loads from.  Rows past the end are clamped to the last row (computed,
// never stored).
__device__ __forceinline__ void set_tile(int64_t tile, int64_t gx, int wave, int lr,
                                         const uint32_t *R, int64_t nR, const uint32_t *Q,
                                         int64_t nQ, int64_t n_words, int64_t &rb, int64_t &qb,
                                         const uint32_t *(&rrow)[kRF],
                                         const uint32_t *(&qrow)[kQF]) {
    const int64_t ty = tile / gx, tx = tile - ty * gx;
    rb = ty * kRefTile + (wave >> 1) * (16 * kRF);
    qb = tx * kQueryTile + (wave & 1) * (16 * kQF);
#pragma unroll
    for (int f = 0; f < kRF; ++f) {
        const int64_t r = rb + 16 * f + lr;
        rrow[f] = R + (r < nR ? r : nR - 1) * n_words;
    }
#pragma unroll
    for (int f = 0; f < kQF; ++f) {
        const int64_t q = qb + 16 * f + lr;
        qrow[f] = Q + (q < nQ ? q : nQ - 1) * n_words;
    }
}

// Measures of one wave tile from its accumulators and lane popcounts.
template <int kMeasure, bool kShort>
__device__ __forceinline__ void fp_epilogue(const v4f (&acc)[kQF][kRF], uint32_t (&pr)[kRF],
                                            uint32_t (&pq)[kQF], int64_t rb, int64_t qb,
                                            int lr, int grp, int64_t nR, int64_t nQ,
                                            const uint8_t *__restrict__ ref_valid,
                                            const uint8_t *__restrict__ query_valid,
                                            float empty_score, float *__restrict__ out,
                                            int32_t *__restrict__ inter_out, int64_t ld,
                                            int vec_store) {
    // Row popcounts: sum the 4 lane groups; lane (lr, *) then holds the count
    // of row 16 f + lr of each fragment.
#pragma unroll
    for (int f = 0; f < kRF; ++f) {
        pr[f] += __shfl_xor(pr[f], 16);
        pr[f] += __shfl_xor(pr[f], 32);
    }
#pragma unroll
    for (int f = 0; f < kQF; ++f) {
        pq[f] += __shfl_xor(pq[f], 16);
        pq[f] += __shfl_xor(pq[f], 32);
    }
    // Accumulator layout (16 x 16 f32): lane (lr, grp) element e holds
    // query 4 grp + e (operand A row), reference lr (operand B row).
    float qcount[kQF][4];
#pragma unroll
    for (int f = 0; f < kQF; ++f)
#pragma unroll
        for (int e = 0; e < 4; ++e) qcount[f][e] = count_floor(__shfl(pq[f], 4 * grp + e));
    const bool flagged = ref_valid || query_valid;

#pragma unroll
    for (int rf = 0; rf < kRF; ++rf) {
        const int64_t r = rb + 16 * rf + lr;
        if (r >= nR) continue;
        const float a = count_floor(pr[rf]);
        const bool rv = !ref_valid || ref_valid[r];
#pragma unroll
        for (int qf = 0; qf < kQF; ++qf) {
            const int64_t q = qb + 16 * qf + 4 * grp;
            v4f v;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                v[e] = measure_value<kMeasure, kShort>(acc[qf][rf][e], a, qcount[qf][e]);
            if (flagged) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const bool qv = !query_valid || (q + e < nQ && query_valid[q + e]);
                    v[e] = (rv & qv) ? v[e] : empty_score;
                }
            }
            float *o = out + r * ld + q;
            if (vec_store && q + 3 < nQ) {
                __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(o));
                if (inter_out) {
                    const v4i iv{(int)acc[qf][rf][0], (int)acc[qf][rf][1], (int)acc[qf][rf][2],
                                 (int)acc[qf][rf][3]};
                    __builtin_nontemporal_store(iv, reinterpret_cast<v4i *>(inter_out + r * ld + q));
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (q + e >= nQ) break;
                    o[e] = v[e];
                    if (inter_out) inter_out[r * ld + q + e] = (int)acc[qf][rf][e];
                }
            }
        }
    }
}

// Persistent: workgroup b handles block tiles b, b + gridDim.x, ... (gx
// tiles per row of tiles).  The first step of the next tile is loaded during
// the last step of the current one, so its latency hides behind the MFMAs
// and the epilogue.
template <bool kFull, int kMeasure, bool kShort>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_fp_gemm(const uint32_t *__restrict__ R, int64_t nR, const uint32_t *__restrict__ Q, int64_t nQ,
          int64_t n_words, int64_t gx, int64_t n_tiles, const uint8_t *__restrict__ ref_valid,
          const uint8_t *__restrict__ query_valid, float empty_score,
          float *__restrict__ out, int32_t *__restrict__ inter_out, int64_t ld, int vec_store) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, grp = lane >> 4;
    int64_t tile = blockIdx.x;
    int64_t rb, qb, rb_next = 0, qb_next = 0;
    const uint32_t *rrow[kRF], *qrow[kQF];
    set_tile(tile, gx, wave, lr, R, nR, Q, nQ, n_words, rb, qb, rrow, qrow);

    const int steps = (int)((n_words + kStepWords - 1) / kStepWords);
    v4u r0[kRF], q0[kQF], r1[kRF], q1[kQF];
    load_step<kFull>(rrow, qrow, 4 * grp, n_words, r0, q0);
    for (;;) {
        v4f acc[kQF][kRF];
#pragma unroll
        for (int i = 0; i < kQF; ++i)
#pragma unroll
            for (int j = 0; j < kRF; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        uint32_t pr[kRF], pq[kQF];
#pragma unroll
        for (int f = 0; f < kRF; ++f) pr[f] = 0u;
#pragma unroll
        for (int f = 0; f < kQF; ++f) pq[f] = 0u;
        const int64_t next = tile + gridDim.x;
        // Two steps per iteration so the prefetch buffers swap roles without
        // copies.  The full-word path has an even step count (checked on the
        // host); the bounded path reads zero words past the end (no bits, no
        // effect).  The last iteration prefetches the next tile's first step.
        for (int s = 0; s < steps; s += 2) {
            const int64_t w1 = (int64_t)(s + 1) * kStepWords + 4 * grp;
            load_step<kFull>(rrow, qrow, w1, n_words, r1, q1);
            fp_step(r0, q0, acc, pr, pq);
            // one load site for the step after (the tile's next, or the next
            // tile's first): two would make the compiler copy the buffers
            const bool in_tile = s + 2 < steps;
            if (!in_tile && next < n_tiles)
                set_tile(next, gx, wave, lr, R, nR, Q, nQ, n_words, rb_next, qb_next, rrow,
                         qrow);
            if (in_tile || next < n_tiles)
                load_step<kFull>(rrow, qrow, in_tile ? w1 + kStepWords : 4 * grp, n_words, r0,
                                 q0);
            fp_step(r1, q1, acc, pr, pq);
        }
        fp_epilogue<kMeasure, kShort>(acc, pr, pq, rb, qb, lr, grp, nR, nQ, ref_valid, query_valid,
                              empty_score, out, inter_out, ld, vec_store);
        if (next >= n_tiles) break;
        tile = next;
        rb = rb_next;
        qb = qb_next;
    }
}

// Compute units of the current device (cached per device; 256 if unknown).
int device_cu_count() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v < 1)
        v = 256;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_fingerprint_workspace_size(int64_t n_refs, int64_t n_queries,
                                                   int64_t n_words, size_t *bytes) {
    if (n_refs < 0 || n_queries < 0 || n_words < 0 || !bytes)
        return fail(CS_ERR_INVALID, "invalid fingerprint shape");
    *bytes = 0;  // the kernel reads the packed rows directly
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
    (void)workspace;
    (void)workspace_bytes;
    if (measure != CS_JACCARD && measure != CS_DICE && measure != CS_COSINE)
        return fail(CS_ERR_INVALID, "unknown measure %d", measure);
    if (n_refs < 0 || n_queries < 0 || n_words < 0)
        return fail(CS_ERR_INVALID, "negative fingerprint shape");
    if (n_refs == 0 || n_queries == 0) return CS_OK;
    if (n_words == 0 || !ref_words || !query_words || !out)
        return fail(CS_ERR_INVALID, "fingerprint pointers must not be NULL and n_words > 0");
    if (ld < n_queries) return fail(CS_ERR_INVALID, "ld < number of queries");
    if (n_words * 32 > kMaxBits)
        return fail(CS_ERR_UNSUPPORTED, "fingerprints longer than %lld bits",
                    (long long)kMaxBits);
    const int64_t gx = (n_queries + kQueryTile - 1) / kQueryTile;
    const int64_t gy = (n_refs + kRefTile - 1) / kRefTile;
    const int64_t n_tiles = gx * gy;
    auto aligned16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const bool full =
        n_words % (2 * kStepWords) == 0 && aligned16(ref_words) && aligned16(query_words);
    const int vec = ld % 4 == 0 && aligned16(out) && (!intersections || aligned16(intersections));
    // Two resident workgroups per CU (256 threads, two waves per SIMD).
    const int64_t grid = std::min<int64_t>(n_tiles, 2 * (int64_t)device_cu_count());
    hipStream_t hs = static_cast<hipStream_t>(stream);
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(kThreads), 0, hs, ref_words,
                           n_refs, query_words, n_queries, n_words, gx, n_tiles, ref_valid,
                           query_valid, empty_score, out, intersections, ld, vec);
    };
    // the short division is exact for denominators up to 2 * 4096 (dice)
    const bool shrt = n_words * 32 <= kShortDivBits;
    if (full) {
        if (measure == CS_JACCARD)
            shrt ? launch(k_fp_gemm<true, CS_JACCARD, true>) : launch(k_fp_gemm<true, CS_JACCARD, false>);
        else if (measure == CS_DICE)
            shrt ? launch(k_fp_gemm<true, CS_DICE, true>) : launch(k_fp_gemm<true, CS_DICE, false>);
        else
            launch(k_fp_gemm<true, CS_COSINE, false>);
    } else {
        if (measure == CS_JACCARD)
            shrt ? launch(k_fp_gemm<false, CS_JACCARD, true>) : launch(k_fp_gemm<false, CS_JACCARD, false>);
        else if (measure == CS_DICE)
            shrt ? launch(k_fp_gemm<false, CS_DICE, true>) : launch(k_fp_gemm<false, CS_DICE, false>);
        else
            launch(k_fp_gemm<false, CS_COSINE, false>);
    }
    CS_LAUNCH_CHECK("k_fp_gemm");
    return CS_OK;
}
