// cg_lane.h -- per-pair CosineGreedy logic shared by the device kernels and
// the host-side unit harness (tests/native/tile_harness.cpp).
//
// Everything here is __host__ __device__ and free of wave intrinsics, so the
// exact code a GPU lane runs to settle one (reference, query) pair can also be
// run on the CPU against the oracle.
//
// Semantics (matchms CosineGreedy.pair as the reference calls it,
// cudams/similarity/spectrum_similarity_functions.py:306-317):
//   candidate (i, j)  iff  lb_i <= mz_q[j] + shift <= hb_i  with
//                    lb_i = mz_r[i] - tol, hb_i = mz_r[i] + tol in float64
//                    (find_matches bounds; the reference GPU scan is :123-149)
//   value            v = ((mz_pow_r[i] * int_pow_r[i]) * mz_pow_q[j]) * int_pow_q[j]
//   order            v descending, ties: later-collected first.  Collection
//                    order is (i, j) ascending, so the total order is
//                    (v desc, i desc, j desc) -- stable ascending sort then
//                    reverse (:314), which the reference GPU mergesort
//                    reproduces by taking the right run on ties (:178-182).
//   greedy           accept iff i and j are both unused (:208-218)
//   score            sum of accepted v (in order) / (norm_r * norm_q) (:261)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CS_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define CS_HD inline
#endif

namespace cs {

// Sort key of one candidate: (v, packed (i << 16) | j), compared
// lexicographically.  Peak indices fit 16 bits (max_peaks <= 65535), so the
// packed integer orders (i desc, j desc) exactly as the collection order does.
// Written with non-short-circuit operators so that the device code is
// branch-free (two compares and a select instead of exec-mask branches).
CS_HD bool key_greater(double va, uint32_t ia, double vb, uint32_t ib) {
    return (va > vb) | ((va == vb) & (ia > ib));
}

CS_HD uint32_t pack_ij(int i, int j) { return (uint32_t(i) << 16) | uint32_t(j); }

// float64 window test exactly as matchms find_matches writes it.
CS_HD bool exact_window(double mz_r, double mz_q, double tol, double shift) {
    const double x = mz_q + shift;
    return !(x > mz_r + tol) & !(x < mz_r - tol);
}

// The same test for shift == 0: mz_q + 0.0 equals mz_q for every mz_q but
// -0.0 (which becomes +0.0), and the comparisons do not tell the two apart,
// so the addition can be left out.
CS_HD bool exact_window_noshift(double mz_r, double mz_q, double tol) {
    return !(mz_q > mz_r + tol) & !(mz_q < mz_r - tol);
}

// Up to K candidates kept sorted by key_greater (descending) in registers.
// All indexing is compile-time after unrolling, so the arrays stay in VGPRs.
// The unrolled steps past the wave's bound are left by an early exit for
// K <= 8; for larger K a loop with an exit in the middle is not unrolled
// (the arrays would go to scratch memory), so those steps are skipped one by
// one instead.
template <int K>
constexpr bool kLaneEarlyExit = K <= 8;

template <int K>
struct LaneCandidates {
    double v[K];
    uint32_t ij[K];
    int n;

    CS_HD void reset() { n = 0; }

    // Insert (cv, cij); caller guarantees n < K and n < bound <= K (bound as
    // in greedy()).
    CS_HD void insert(double cv, uint32_t cij, int bound = K) {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            if (s >= bound) {
                if constexpr (kLaneEarlyExit<K>)
                    break;
                else
                    continue;
            }
            const bool take = (s >= n) | key_greater(cv, cij, v[s], ij[s]);
            const double tv = take ? v[s] : cv;
            const uint32_t tij = take ? ij[s] : cij;
            v[s] = take ? cv : v[s];
            ij[s] = take ? cij : ij[s];
            cv = tv;
            cij = tij;
        }
        ++n;
    }

    // Insert the candidate of rank s when the s before it are already held
    // (n == s): it is compared with positions 0 .. s - 1 only and the last
    // one moves to position s.  With s a constant after unrolling, inserting
    // B candidates costs B (B - 1) / 2 compare-exchanges instead of B * B.
    CS_HD void insert_rank(int s, double cv, uint32_t cij) {
#pragma unroll
        for (int p = 0; p < K; ++p) {
            if (p > s) break;
            const bool take = (p == s) | key_greater(cv, cij, v[p], ij[p]);
            const double tv = take ? v[p] : cv;
            const uint32_t tij = take ? ij[p] : cij;
            v[p] = take ? cv : v[p];
            ij[p] = take ? cij : ij[p];
            cv = tv;
            cij = tij;
        }
        n = s + 1;
    }

    // Greedy unique assignment over the sorted candidates.  `bound` >= n may
    // be smaller than K; on the device it is the same for the whole wave, so
    // the unrolled steps past it are skipped by a uniform branch.
    CS_HD void greedy(double &score, int &matches, int bound = K) const {
        double s = 0.0;
        int m = 0;
        bool acc[K];
#pragma unroll
        for (int a = 0; a < K; ++a) {
            if (a >= bound) {
                if constexpr (kLaneEarlyExit<K>)
                    break;
                else
                    continue;
            }
            bool ok = a < n;
            const uint32_t ia = ij[a] >> 16, ja = ij[a] & 0xffffu;
#pragma unroll
            for (int b = 0; b < a; ++b) {
                const bool clash = ((ij[b] >> 16) == ia) | ((ij[b] & 0xffffu) == ja);
                ok = ok & !(acc[b] & clash);
            }
            acc[a] = ok;
            if (ok) {
                s += v[a];
                ++m;
            }
        }
        score = s;
        matches = m;
    }
};

// Coarse float32 tolerance covering every rounding error of the float32 scan
// relative to the float64 test: m/z rounding of both peaks, the shift and its
// float32 sum, the float32 difference, and the float64 rounding of
// mz_r +- tol and mz_q + shift.  Each is at most 2^-24 of (max|mz| + |shift| +
// tol); 8 of them bound the total, and the result is rounded up.
inline float coarse_tolerance(double tol, double shift, double max_abs_mz) {
    const double m = (max_abs_mz < 0 ? -max_abs_mz : max_abs_mz) + (shift < 0 ? -shift : shift) + tol;
    const double e = 8.0 * m / 16777216.0;
    const double t = tol + e;
    float f = (float)t;
    if ((double)f < t) f = __builtin_nextafterf(f, __builtin_huge_valf());
    return __builtin_nextafterf(f, __builtin_huge_valf());
}

}  // namespace cs
