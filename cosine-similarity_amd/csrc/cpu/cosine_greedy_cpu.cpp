// cosine_greedy_cpu.cpp -- all-pairs CosineGreedy on the host cores (OpenMP),
// the engine behind CPUParallelCosineGreedy.
//
// Replaces the numba prange kernel of cpu_parallel_cosine_greedy_kernel
// (cudams/similarity/spectrum_similarity_functions.py:281-320), which runs
// matchms collect_peak_pairs -> mergesort[::-1] -> score_best_matches per
// pair, for the CPUParallelCosineGreedy class (CPUParallelCosineGreedy.py:
// 19-108).  Results are matchms CosineGreedy's for every pair (the reference
// kernel's match_limit never truncates: matchms collects every candidate).
//
// The work is organised like the GPU path, with the same shared code
// (cg_tile.h, cg_lane.h): 64 consecutive queries form a tile whose peaks are
// merged into one list sorted by (m/z, lane, peak) with a bucket directory.
// A reference peak finds its window in the whole tile with one directory
// lookup and a short walk, instead of one merge scan per (reference, query)
// pair; most pairs have no candidate and cost nothing beyond that probe.
// A work item is one tile against a block of references, so the tile stays
// in the core's cache while the block is probed.  Candidates are re-tested
// with matchms's float64 window, ordered by (v desc, i desc, j desc) and
// accepted greedily; the score is the float64 sum in acceptance order divided
// by the two norms, rounded to float32 once.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cg_tile.h"
#include "cosine_similarity_amd.h"
#include "cpu_common.h"

namespace {

using cs::kTileBuckets;
using cs::kTileQueries;

constexpr int kRefBlock = 64;     // references per work item
constexpr int kSmallPair = 16;    // candidates settled by insertion sort

struct Tile {
    int64_t q0 = 0;
    int nl = 0;
    float lo = 0.0f, iw = 0.0f;
    std::vector<float> smz;      // [np + 1], +inf pad
    std::vector<uint32_t> who;   // lane << 16 | peak
    std::vector<int32_t> dir;    // [kTileBuckets + 1]
};

struct Cand {
    double v;
    uint32_t ij;
};

inline double exact_mz(const cs_spectra &s, int64_t g) {
    return s.mz_exact ? s.mz_exact[g] : (double)s.mz[g];
}

inline double weight(const cs_spectra &s, int64_t g) {
    return s.mz_pow ? s.mz_pow[g] * s.int_pow[g] : 1.0 * s.int_pow[g];
}

void build_tile(const cs_spectra &Q, float shift32, int64_t q0, Tile &t) {
    t.q0 = q0;
    t.nl = (int)std::min<int64_t>(kTileQueries, Q.n_spectra - q0);
    struct E {
        float x;
        uint32_t who;
    };
    std::vector<E> e;
    const int64_t base = Q.offsets[q0];
    e.reserve((size_t)(Q.offsets[q0 + t.nl] - base));
    for (int l = 0; l < t.nl; ++l)
        for (int64_t g = Q.offsets[q0 + l]; g < Q.offsets[q0 + l + 1]; ++g)
            e.push_back({Q.mz[g] + shift32, cs::tile_who(l, (int)(g - Q.offsets[q0 + l]))});
    // (value, lane, peak): who orders lane then peak
    std::sort(e.begin(), e.end(), [](const E &a, const E &b) {
        return a.x < b.x || (a.x == b.x && a.who < b.who);
    });
    const int np = (int)e.size();
    t.smz.resize(np + 1);
    t.who.resize(np + 1);
    for (int p = 0; p < np; ++p) {
        t.smz[p] = e[p].x;
        t.who[p] = e[p].who;
    }
    t.smz[np] = __builtin_huge_valf();
    t.who[np] = 0;
    t.lo = np ? t.smz[0] : 0.0f;
    t.iw = cs::tile_inv_width(t.lo, np ? t.smz[np - 1] : 0.0f);
    t.dir.assign(kTileBuckets + 1, np);
    int b = 0;
    for (int p = 0; p < np && b <= kTileBuckets; ++p) {
        const int bp = cs::tile_bucket(t.smz[p], t.lo, t.iw);
        while (b <= bp) t.dir[b++] = p;
    }
}

// Per-thread state of one work item: the candidates of each query lane, in
// collection order, and the peak-use flags of the greedy.
struct Worker {
    std::vector<Cand> lane[kTileQueries];
    std::vector<unsigned char> used_r, used_q;
    std::vector<uint32_t> accepted;
};

// Greedy over one pair's candidates (collection order on entry).
void settle(std::vector<Cand> &c, Worker &w, double &score, int &matches) {
    const int n = (int)c.size();
    auto greater = [](const Cand &a, const Cand &b) { return cs::key_greater(a.v, a.ij, b.v, b.ij); };
    double s = 0.0;
    int m = 0;
    if (n <= kSmallPair) {
        // insertion sort, then accept a candidate unless an accepted one
        // before it shares a peak
        for (int x = 1; x < n; ++x) {
            const Cand k = c[x];
            int y = x - 1;
            while (y >= 0 && greater(k, c[y])) {
                c[y + 1] = c[y];
                --y;
            }
            c[y + 1] = k;
        }
        uint32_t acc[kSmallPair];
        for (int x = 0; x < n; ++x) {
            const uint32_t i = c[x].ij >> 16, j = c[x].ij & 0xffffu;
            bool ok = true;
            for (int y = 0; y < m; ++y)
                ok &= ((acc[y] >> 16) != i) & ((acc[y] & 0xffffu) != j);
            if (ok) {
                acc[m++] = c[x].ij;
                s += c[x].v;
            }
        }
    } else {
        std::sort(c.begin(), c.end(), greater);
        w.accepted.clear();
        for (int x = 0; x < n; ++x) {
            const uint32_t i = c[x].ij >> 16, j = c[x].ij & 0xffffu;
            if (!w.used_r[i] && !w.used_q[j]) {
                w.used_r[i] = 1;
                w.used_q[j] = 1;
                w.accepted.push_back(c[x].ij);
                s += c[x].v;
                ++m;
            }
        }
        for (uint32_t ij : w.accepted) {
            w.used_r[ij >> 16] = 0;
            w.used_q[ij & 0xffffu] = 0;
        }
    }
    score = s;
    matches = m;
}

void run_item(const cs_spectra &R, const cs_spectra &Q, const Tile &t, int64_t r0, int64_t r1,
              double tol, double shift, float tolc, float *score, int32_t *matches, int64_t ld,
              Worker &w) {
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t a = R.offsets[r], nr = R.offsets[r + 1] - a;
        for (int l = 0; l < t.nl; ++l) w.lane[l].clear();
        for (int64_t i = 0; i < nr; ++i) {
            const float ri = R.mz[a + i];
            const float x0 = ri - tolc, x1 = ri + tolc;
            const double mr = exact_mz(R, a + i);
            const double wr = weight(R, a + i);
            cs::tile_probe(
                x0, x1, t.lo, t.iw, [&](int p) { return t.smz[p]; },
                [&](int b) { return t.dir[b]; },
                [&](int p) {
                    const int l = (int)(t.who[p] >> 16), j = (int)(t.who[p] & 0xffffu);
                    const int64_t g = Q.offsets[t.q0 + l] + j;
                    if (!cs::exact_window(mr, exact_mz(Q, g), tol, shift)) return;
                    // ((mz_r^p * I_r^q) * mz_q^p) * I_q^q, left to right
                    const double v =
                        (wr * (Q.mz_pow ? Q.mz_pow[g] : 1.0)) * Q.int_pow[g];
                    w.lane[l].push_back({v, cs::pack_ij((int)i, j)});
                });
        }
        for (int l = 0; l < t.nl; ++l) {
            const int64_t q = t.q0 + l;
            double s = 0.0;
            int m = 0;
            if (!w.lane[l].empty()) settle(w.lane[l], w, s, m);
            const int64_t o = r * ld + q;
            score[o] = m ? (float)(s / (R.norm[r] * Q.norm[q])) : 0.0f;
            matches[o] = m;
        }
    }
}

}  // namespace

extern "C" cs_status cs_cpu_cosine_greedy(const cs_spectra *refs, const cs_spectra *queries,
                                          double tolerance, double shift, float *score,
                                          int32_t *matches, int64_t ld, int n_threads) {
    cs_status st;
    if ((st = cs::cpu::check_host_spectra(refs, "refs")) != CS_OK) return st;
    if ((st = cs::cpu::check_host_spectra(queries, "queries")) != CS_OK) return st;
    if (!(tolerance >= 0.0) || tolerance == __builtin_huge_val())
        return cs::cpu::fail(CS_ERR_INVALID, "tolerance must be finite and >= 0");
    if (!(shift == shift) || shift == __builtin_huge_val() || shift == -__builtin_huge_val())
        return cs::cpu::fail(CS_ERR_INVALID, "shift must be finite");
    const cs_spectra &R = *refs, &Q = *queries;
    if (R.n_spectra == 0 || Q.n_spectra == 0) return CS_OK;
    if (!score || !matches) return cs::cpu::fail(CS_ERR_INVALID, "output pointer is NULL");
    if (ld < Q.n_spectra)
        return cs::cpu::fail(CS_ERR_INVALID, "ld %lld < number of queries %lld", (long long)ld,
                             (long long)Q.n_spectra);
    const float tolc = cs::coarse_tolerance(tolerance, shift,
                                            std::max(R.max_abs_mz, Q.max_abs_mz));
    const float shift32 = (float)shift;
    const int64_t ntiles = (Q.n_spectra + kTileQueries - 1) / kTileQueries;
    const int64_t nblocks = (R.n_spectra + kRefBlock - 1) / kRefBlock;
    int nt = 1;
#ifdef _OPENMP
    nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
#endif
    std::vector<Tile> tiles((size_t)ntiles);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t t = 0; t < ntiles; ++t) build_tile(Q, shift32, t * kTileQueries, tiles[t]);
#pragma omp parallel num_threads(nt)
    {
        Worker w;
        w.used_r.assign((size_t)R.max_peaks + 1, 0);
        w.used_q.assign((size_t)Q.max_peaks + 1, 0);
        // tile-major items: consecutive items of one thread reuse the tile
#pragma omp for schedule(dynamic, 1)
        for (int64_t item = 0; item < ntiles * nblocks; ++item) {
            const int64_t t = item / nblocks, b = item % nblocks;
            const int64_t r0 = b * kRefBlock, r1 = std::min(R.n_spectra, r0 + kRefBlock);
            run_item(R, Q, tiles[t], r0, r1, tolerance, shift, tolc, score, matches, ld, w);
        }
    }
    return CS_OK;
}
