// cosine_greedy.hip -- all-pairs CosineGreedy on gfx950.
//
// Replaces the Numba-CUDA kernel of cudams/similarity/spectrum_similarity_
// functions.py:39-262 and the per-tile host loop of CudaCosineGreedy.py:149-185.
//
// Pipeline (all asynchronous on the caller's stream, no host sync):
//   1. k_fill            zero counters / histogram (a kernel, not a memset node,
//                        so the sequence can sit in a captured graph)
//   2. k_len_hist, k_scan_block, k_len_scatter
//                        counting sort of the queries by peak count
//   3. k_tile_len, k_scan_block, k_tile_build
//                        groups of 64 length-sorted queries ("query tiles"),
//                        stored interleaved [peak][lane] with a +inf pad row,
//                        so lane l reads its own query with no LDS bank
//                        conflicts whatever peak each lane is at
//   4. k_light<K>        one lane per (reference, query) pair: float32 merge
//                        scan with an exact float64 test on coarse hits; up to
//                        K candidates are sorted and settled in registers
//   5. k_heavy           one wave per pair that had more than K candidates:
//                        parallel window enumeration, LDS bitonic sort by the
//                        (v desc, i desc, j desc) key, wave-parallel greedy,
//                        in rounds of at most T candidates, so any candidate
//                        count is handled exactly in bounded LDS.
#include <limits.h>

#include "cs_internal.h"
#include "cg_lane.h"

namespace cs {
namespace {

constexpr int kWave = 64;
constexpr int kLightThreads = 256;
constexpr int kRefsPerBlock = 32;     // references one light workgroup walks
constexpr int kTileLdsCap = 160;      // query peaks (incl. pad) per lane staged in LDS
constexpr int kDefaultK = 8;
constexpr int kDefaultRound = 512;
constexpr int kHeavyGrid = 2048;
constexpr int64_t kHeavyQueueCap = int64_t(1) << 20;
constexpr int kMaxPeaks = 65535;

struct SpecView {
    const int64_t *off;
    const float *mz;
    const double *mzx;
    const double *mzp;
    const double *inp;
    const double *norm;
};

SpecView view(const cs_spectra &s) {
    return SpecView{s.offsets, s.mz, s.mz_exact, s.mz_pow, s.int_pow, s.norm};
}

__device__ __forceinline__ double exact_mz(const SpecView &s, int64_t p) {
    return s.mzx ? s.mzx[p] : (double)s.mz[p];
}

// ((mz_r^p * I_r^q) * mz_q^p) * I_q^q, the left-to-right order of matchms.
__device__ __forceinline__ double pair_value(const SpecView &R, int64_t pr, const SpecView &Q,
                                             int64_t pq) {
    const double a = R.mzp ? R.mzp[pr] : 1.0;
    const double c = Q.mzp ? Q.mzp[pq] : 1.0;
    double v = a * R.inp[pr];
    v = v * c;
    v = v * Q.inp[pq];
    return v;
}

// ---------------------------------------------------------------------------
// Prep kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t value) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__device__ __forceinline__ int clamp_len(int64_t n, int maxbin) {
    return n < 0 ? 0 : (n > maxbin ? maxbin : (int)n);
}

__global__ void k_len_hist(const int64_t *off, int64_t n, int maxbin, int32_t *hist) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&hist[clamp_len(off[q + 1] - off[q], maxbin)], 1);
}

// Exclusive scan of n int32 values by one 1024-thread block.
// out64[i] = mult * sum(in[0..i)), out64[n] = mult * total; out32 likewise
// (without the total).  Either output may be null.
__global__ void __launch_bounds__(1024) k_scan_block(const int32_t *in, int64_t n, int64_t mult,
                                                     int64_t *out64, int32_t *out32) {
    __shared__ int64_t part[1024];
    const int tid = threadIdx.x;
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t b = tid * chunk;
    const int64_t e = b + chunk < n ? b + chunk : n;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += in[i];
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
        const int32_t x = in[i];
        if (out64) out64[i] = run * mult;
        if (out32) out32[i] = (int32_t)run;
        run += x;
    }
    if (tid == 1023 && out64) out64[n] = part[1023] * mult;
}

__global__ void k_len_scatter(const int64_t *off, int64_t n, int maxbin, int32_t *cursor,
                              int32_t *order) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int pos = atomicAdd(&cursor[clamp_len(off[q + 1] - off[q], maxbin)], 1);
        order[pos] = (int32_t)q;
    }
}

// tile_len[t] = 1 + max peaks over the tile's queries (the +1 is the +inf pad row).
__global__ void k_tile_len(const int64_t *off, const int32_t *order, int64_t nq, int64_t ntiles,
                           int maxbin, int32_t *tile_len) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        int m = 0;
        for (int l = 0; l < kWave; ++l) {
            const int64_t qi = t * kWave + l;
            if (qi < nq) {
                const int q = order[qi];
                const int len = clamp_len(off[q + 1] - off[q], maxbin);
                m = len > m ? len : m;
            }
        }
        tile_len[t] = m + 1;
    }
}

__global__ void k_tile_build(const float *mz, const int64_t *off, const int32_t *order, int64_t nq,
                             const int64_t *tile_base, const int32_t *tile_len, int64_t ntiles,
                             float shift32, float *tiles) {
    const float inf = __builtin_huge_valf();
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int L = tile_len[t];
        float *dst = tiles + tile_base[t];
        for (int e = threadIdx.x; e < kWave * L; e += blockDim.x) {
            const int k = e >> 6, lane = e & 63;
            const int64_t qi = t * kWave + lane;
            float v = inf;
            if (qi < nq) {
                const int q = order[qi];
                const int64_t n = off[q + 1] - off[q];
                if (k < n && k < L - 1) v = mz[off[q] + k] + shift32;
            }
            dst[e] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Light kernel: one lane per pair
// ---------------------------------------------------------------------------

template <int K>
__global__ void __launch_bounds__(kLightThreads)
k_light(SpecView R, int64_t nR, SpecView Q, int64_t nQ, const int32_t *__restrict__ order,
        const int64_t *__restrict__ tile_base, const int32_t *__restrict__ tile_len,
        const float *__restrict__ tiles, float tolc, double tol, double shift,
        int64_t match_limit, float *__restrict__ score, int32_t *__restrict__ matches,
        uint8_t *__restrict__ overflow, int64_t ld, unsigned long long *__restrict__ heavy_queue,
        int64_t heavy_cap, unsigned long long *__restrict__ counters) {
    extern __shared__ float lds_tile[];
    __shared__ int next_ref;

    const int64_t rb = blockIdx.x;
    const int64_t t = blockIdx.y;
    const int L = tile_len[t];
    const float *gt = tiles + tile_base[t];
    const bool in_lds = L <= kTileLdsCap;
    if (in_lds)
        for (int e = threadIdx.x; e < kWave * L; e += kLightThreads) lds_tile[e] = gt[e];
    if (threadIdx.x == 0) next_ref = 0;
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int64_t qi = t * kWave + lane;
    const bool valid = qi < nQ;
    const int q = valid ? order[qi] : 0;
    const int64_t qoff = Q.off[q];
    int nq = valid ? (int)(Q.off[q + 1] - qoff) : 0;
    nq = nq < L - 1 ? nq : L - 1;  // the tile holds L - 1 peaks + the pad row
    const double qnorm = Q.norm[q];
    const int64_t r0 = rb * kRefsPerBlock;
    const int64_t r1 = r0 + kRefsPerBlock < nR ? r0 + kRefsPerBlock : nR;

    for (;;) {
        int rr = 0;
        if (lane == 0) rr = atomicAdd(&next_ref, 1);
        rr = __shfl(rr, 0);
        const int64_t r = r0 + rr;
        if (r >= r1) break;
        const int64_t roff = R.off[r];
        const int nr = (int)(R.off[r + 1] - roff);
        const float *rmz = R.mz + roff;

        LaneCandidates<K> cand;
        auto ref_mz = [&](int k) { return rmz[k]; };
        auto exact = [&](int k, int j) {
            return exact_window(exact_mz(R, roff + k), exact_mz(Q, qoff + j), tol, shift);
        };
        auto value = [&](int k, int j) { return pair_value(R, roff + k, Q, qoff + j); };
        LaneResult res;
        if (in_lds) {
            auto qry = [&](int j) { return lds_tile[j * kWave + lane]; };
            res = scan_pair<K>(nr, nq, tolc, ref_mz, qry, exact, value, cand);
        } else {
            auto qry = [&](int j) { return gt[(int64_t)j * kWave + lane]; };
            res = scan_pair<K>(nr, nq, tolc, ref_mz, qry, exact, value, cand);
        }
        if (valid) {
            const int64_t o = r * ld + q;
            if (res.heavy) {
                const unsigned long long slot = atomicAdd(&counters[0], 1ull);
                if ((int64_t)slot < heavy_cap)
                    heavy_queue[slot] = ((unsigned long long)r << 32) | (unsigned int)q;
                score[o] = 0.0f;
                matches[o] = -1;
                overflow[o] = 0;
            } else {
                score[o] = res.candidates ? (float)(res.score / (R.norm[r] * qnorm)) : 0.0f;
                matches[o] = res.matches;
                overflow[o] = (uint8_t)(res.candidates >= match_limit);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Heavy kernel: one wave per pair with more than K candidates
// ---------------------------------------------------------------------------

struct HeavyArgs {
    SpecView R, Q;
    int64_t nR, nQ;
    float tolc, shift32;
    double tol, shift;
    int64_t match_limit;
    int round_cap;     // T
    int buf_cap;       // next_pow2(2T + 64) entries
    int r_words, q_words;
    float *score;
    int32_t *matches;
    uint8_t *overflow;
    int64_t ld;
    const unsigned long long *heavy_queue;
    int64_t heavy_cap;
    unsigned long long *counters;
};

__device__ __forceinline__ bool bit_set(const uint32_t *b, int x) {
    return (b[x >> 5] >> (x & 31)) & 1u;
}

// Pad [n, next_pow2(n)) with -inf keys and sort the first next_pow2(n)
// entries descending by key_greater.  One wave, LDS.
__device__ void sort_desc(double *bv, uint32_t *bij, int n) {
    const int lane = threadIdx.x;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int e = n + lane; e < n2; e += kWave) {
        bv[e] = -__builtin_huge_val();
        bij[e] = 0u;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (n2 >> 1); t += kWave) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool desc = (lo & size) == 0;
                const double va = bv[lo], vb = bv[hi];
                const uint32_t ia = bij[lo], ib = bij[hi];
                const bool sw = desc ? key_greater(vb, ib, va, ia) : key_greater(va, ia, vb, ib);
                if (sw) {
                    bv[lo] = vb;
                    bv[hi] = va;
                    bij[lo] = ib;
                    bij[hi] = ia;
                }
            }
            __syncthreads();
        }
    }
}

// Greedy over sorted entries [0, n): chunks of 64 in order; inside a chunk the
// first live candidate is accepted and every later one sharing its i or j is
// dropped, repeatedly, which is exactly the sequential greedy.
__device__ void greedy_sorted(const double *bv, const uint32_t *bij, int n, uint32_t *used_r,
                              uint32_t *used_q, double &s, int &m) {
    const int lane = threadIdx.x;
    for (int base = 0; base < n; base += kWave) {
        const int idx = base + lane;
        const bool valid = idx < n;
        const uint32_t ij = valid ? bij[idx] : 0u;
        const double v = valid ? bv[idx] : 0.0;
        const int i = (int)(ij >> 16), j = (int)(ij & 0xffffu);
        bool alive = valid && !bit_set(used_r, i) && !bit_set(used_q, j);
        unsigned long long mask = __ballot(alive);
        while (mask) {
            const int f = __builtin_ctzll(mask);
            const int fi = __shfl(i, f), fj = __shfl(j, f);
            const double fv = __shfl(v, f);
            s += fv;
            ++m;
            if (lane == 0) {
                used_r[fi >> 5] |= 1u << (fi & 31);
                used_q[fj >> 5] |= 1u << (fj & 31);
            }
            alive = alive && lane != f && i != fi && j != fj;
            mask = __ballot(alive);
        }
        __syncthreads();
    }
}

__device__ void heavy_pair(const HeavyArgs &a, int64_t r, int64_t q, uint32_t *used_r,
                           uint32_t *used_q, double *bv, uint32_t *bij) {
    const int lane = threadIdx.x;
    const int64_t roff = a.R.off[r], qoff = a.Q.off[q];
    const int nr = (int)(a.R.off[r + 1] - roff), nq = (int)(a.Q.off[q + 1] - qoff);
    const int nrw = (nr + 31) >> 5, nqw = (nq + 31) >> 5;
    const int64_t o = r * a.ld + q;
    if (nrw > a.r_words || nqw > a.q_words) {  // caller's max_peaks was wrong
        if (lane == 0) {
            a.score[o] = __builtin_nanf("");
            a.matches[o] = -2;
            a.overflow[o] = 1;
        }
        return;
    }
    for (int w = lane; w < nrw; w += kWave) used_r[w] = 0u;
    for (int w = lane; w < nqw; w += kWave) used_q[w] = 0u;
    __syncthreads();

    const float *rmz = a.R.mz + roff;
    const float *qmz = a.Q.mz + qoff;
    const int T = a.round_cap;
    bool have_up = false;
    double up_v = 0.0;
    uint32_t up_ij = 0u;
    long long total = 0;
    double s = 0.0;
    int m = 0;

    for (int round = 0;; ++round) {
        int bn = 0;
        bool compacted = false;
        double cut_v = 0.0;
        uint32_t cut_ij = 0u;
        for (int ib = 0; ib < nr; ib += kWave) {
            const int i = ib + lane;
            bool act = i < nr;
            const float ri = act ? rmz[i] : 0.0f;
            int j = nq;
            if (act) {  // first query peak not above the window of reference peak i
                int lo = 0, hi = nq;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    const float d = ri - (qmz[mid] + a.shift32);
                    if (d > a.tolc) lo = mid + 1;
                    else hi = mid;
                }
                j = lo;
            }
            act = act && j < nq;
            while (__ballot(act)) {
                bool push = false;
                double v = 0.0;
                uint32_t ij = 0u;
                if (act) {
                    const float d = ri - (qmz[j] + a.shift32);
                    if (d < -a.tolc) {
                        act = false;
                    } else {
                        if (exact_window(exact_mz(a.R, roff + i), exact_mz(a.Q, qoff + j), a.tol,
                                         a.shift)) {
                            if (round == 0) ++total;
                            ij = pack_ij(i, j);
                            if (!bit_set(used_r, i) && !bit_set(used_q, j)) {
                                v = pair_value(a.R, roff + i, a.Q, qoff + j);
                                push = (!have_up || key_greater(up_v, up_ij, v, ij)) &&
                                       (!compacted || key_greater(v, ij, cut_v, cut_ij));
                            }
                        }
                        ++j;
                        act = j < nq;
                    }
                }
                // append (bn <= 2T before, so bn + 64 fits the buffer) ...
                const unsigned long long mask = __ballot(push);
                if (push) {
                    const int pos = bn + __popcll(mask & ((1ull << lane) - 1ull));
                    bv[pos] = v;
                    bij[pos] = ij;
                }
                bn += __popcll(mask);
                // ... then keep the best T and raise the cut to the T-th key
                if (bn > 2 * T) {
                    sort_desc(bv, bij, bn);
                    bn = T;
                    cut_v = bv[T - 1];
                    cut_ij = bij[T - 1];
                    compacted = true;
                    __syncthreads();
                }
            }
        }
        __syncthreads();
        sort_desc(bv, bij, bn);
        const int take = compacted ? T : bn;
        const double last_v = take > 0 ? bv[take - 1] : 0.0;
        const uint32_t last_ij = take > 0 ? bij[take - 1] : 0u;
        greedy_sorted(bv, bij, take, used_r, used_q, s, m);
        if (!compacted) break;
        have_up = true;
        up_v = last_v;
        up_ij = last_ij;
    }

    for (int off = kWave / 2; off > 0; off >>= 1) total += __shfl_xor(total, off);
    if (lane == 0) {
        a.score[o] = (float)(s / (a.R.norm[r] * a.Q.norm[q]));
        a.matches[o] = m;
        a.overflow[o] = (uint8_t)(total >= a.match_limit);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kWave) k_heavy(HeavyArgs a) {
    extern __shared__ __align__(16) unsigned char heavy_lds[];
    double *bv = reinterpret_cast<double *>(heavy_lds);
    uint32_t *bij = reinterpret_cast<uint32_t *>(bv + a.buf_cap);
    uint32_t *used_r = bij + a.buf_cap;
    uint32_t *used_q = used_r + a.r_words;

    const unsigned long long count = a.counters[0];
    if (count == 0) return;
    const int lane = threadIdx.x;
    if ((int64_t)count <= a.heavy_cap) {
        for (;;) {
            unsigned long long item = 0;
            if (lane == 0) item = atomicAdd(&a.counters[1], 1ull);
            item = __shfl(item, 0);
            if (item >= count) break;
            const unsigned long long rq = a.heavy_queue[item];
            heavy_pair(a, (int64_t)(rq >> 32), (int64_t)(rq & 0xffffffffull), used_r, used_q, bv,
                       bij);
        }
    } else {  // queue overflowed: find the -1 sentinels
        const int64_t total = a.nR * a.nQ;
        for (int64_t c = blockIdx.x; c * kWave < total; c += gridDim.x) {
            const int64_t p = c * kWave + lane;
            int64_t r = 0, q = 0;
            int32_t mv = 0;
            if (p < total) {
                r = p / a.nQ;
                q = p - r * a.nQ;
                mv = a.matches[r * a.ld + q];
            }
            unsigned long long mask = __ballot(p < total && mv == -1);
            while (mask) {
                const int f = __builtin_ctzll(mask);
                const int64_t rf = __shfl(r, f), qf = __shfl(q, f);
                heavy_pair(a, rf, qf, used_r, used_q, bv, bij);
                mask &= mask - 1ull;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct Layout {
    unsigned long long *counters;
    int32_t *hist, *cursor, *order, *tile_len;
    int64_t *tile_base;
    unsigned long long *heavy_queue;
    float *tiles;
    int64_t ntiles, nbins, heavy_cap, tile_elems;
    size_t bytes;
};

Layout layout(void *base, const cs_spectra &R, const cs_spectra &Q) {
    Layout L{};
    Carver c(base);
    L.ntiles = (Q.n_spectra + kWave - 1) / kWave;
    L.nbins = (int64_t)Q.max_peaks + 1;
    L.heavy_cap = R.n_spectra * Q.n_spectra < kHeavyQueueCap ? R.n_spectra * Q.n_spectra
                                                             : kHeavyQueueCap;
    // Tiles are length-sorted, so the longest query of tile t is no longer than
    // the shortest of tile t + 1: sum of tile maxima <= n_peaks / 64 + 2 * max_peaks.
    L.tile_elems = kWave * ((Q.n_peaks + kWave - 1) / kWave + 2 * (int64_t)Q.max_peaks +
                            L.ntiles + 1);
    // query-only buffers first: their offsets do not depend on the references,
    // so a PREPARE done once serves later calls with other references
    L.hist = c.take<int32_t>(L.nbins);
    L.cursor = c.take<int32_t>(L.nbins);
    L.order = c.take<int32_t>(Q.n_spectra);
    L.tile_len = c.take<int32_t>(L.ntiles);
    L.tile_base = c.take<int64_t>(L.ntiles + 1);
    L.tiles = c.take<float>(L.tile_elems);
    L.counters = c.take<unsigned long long>(4);
    L.heavy_queue = c.take<unsigned long long>(L.heavy_cap > 0 ? L.heavy_cap : 1);
    L.bytes = align_up(c.used, 256);
    return L;
}

cs_status check_spectra(const cs_spectra *s, const char *name) {
    if (!s) return fail(CS_ERR_INVALID, "%s is NULL", name);
    if (s->n_spectra < 0 || s->n_peaks < 0)
        return fail(CS_ERR_INVALID, "%s: negative size", name);
    if (s->n_spectra > INT_MAX - kWave)
        return fail(CS_ERR_UNSUPPORTED, "%s: more than 2^31 spectra", name);
    if (s->max_peaks < 0 || s->max_peaks > kMaxPeaks)
        return fail(CS_ERR_UNSUPPORTED, "%s: max_peaks %d outside [0, %d]", name, s->max_peaks,
                    kMaxPeaks);
    if (s->n_spectra > 0 && (!s->offsets || !s->norm))
        return fail(CS_ERR_INVALID, "%s: offsets/norm must not be NULL", name);
    if (s->n_peaks > 0 && (!s->mz || !s->int_pow))
        return fail(CS_ERR_INVALID, "%s: mz/int_pow must not be NULL", name);
    if (!(s->max_abs_mz >= 0.0f) || s->max_abs_mz == __builtin_huge_valf())
        return fail(CS_ERR_INVALID, "%s: max_abs_mz must be finite", name);
    return CS_OK;
}

template <int K>
cs_status launch_light(dim3 grid, hipStream_t st, const SpecView &R, int64_t nR, const SpecView &Q,
                       int64_t nQ, const Layout &L, float tolc, double tol, double shift,
                       int64_t ml, float *score, int32_t *matches, uint8_t *overflow, int64_t ld) {
    const size_t lds = sizeof(float) * kWave * kTileLdsCap;
    hipLaunchKernelGGL(k_light<K>, grid, dim3(kLightThreads), lds, st, R, nR, Q, nQ, L.order,
                       L.tile_base, L.tile_len, L.tiles, tolc, tol, shift, ml, score, matches,
                       overflow, ld, L.heavy_queue, L.heavy_cap, L.counters);
    CS_LAUNCH_CHECK("k_light");
    return CS_OK;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_cosine_greedy_workspace_size(const cs_spectra *refs,
                                                     const cs_spectra *queries, size_t *bytes) {
    cs_status st;
    if ((st = check_spectra(refs, "refs")) != CS_OK) return st;
    if ((st = check_spectra(queries, "queries")) != CS_OK) return st;
    if (!bytes) return fail(CS_ERR_INVALID, "bytes is NULL");
    *bytes = layout(nullptr, *refs, *queries).bytes;
    return CS_OK;
}

extern "C" cs_status cs_cosine_greedy(const cs_spectra *refs, const cs_spectra *queries,
                                      const cs_cosine_params *params, float *score,
                                      int32_t *matches, uint8_t *overflow, int64_t ld,
                                      void *workspace, size_t workspace_bytes, void *stream) {
    cs_status st;
    if ((st = check_spectra(refs, "refs")) != CS_OK) return st;
    if ((st = check_spectra(queries, "queries")) != CS_OK) return st;
    if (!params) return fail(CS_ERR_INVALID, "params is NULL");
    const cs_spectra &Rs = *refs, &Qs = *queries;
    const int64_t nR = Rs.n_spectra, nQ = Qs.n_spectra;
    if (nR == 0 || nQ == 0) return CS_OK;
    if (!score || !matches || !overflow) return fail(CS_ERR_INVALID, "output pointer is NULL");
    if (ld < nQ) return fail(CS_ERR_INVALID, "ld %lld < number of queries %lld", (long long)ld,
                             (long long)nQ);
    const double tol = params->tolerance, shift = params->shift;
    if (!(tol >= 0.0) || tol == __builtin_huge_val())
        return fail(CS_ERR_INVALID, "tolerance must be finite and >= 0");
    if (!(shift == shift) || shift == __builtin_huge_val() || shift == -__builtin_huge_val())
        return fail(CS_ERR_INVALID, "shift must be finite");
    int K = params->light_candidates ? params->light_candidates : kDefaultK;
    if (K != 1 && K != 4 && K != 8 && K != 16)
        return fail(CS_ERR_INVALID, "light_candidates must be 1, 4, 8 or 16 (got %d)", K);
    const int T = params->round_capacity ? params->round_capacity : kDefaultRound;
    if (T < 4 || T > 1024 || (T & (T - 1)))
        return fail(CS_ERR_INVALID, "round_capacity must be a power of two in [4, 1024]");

    const Layout need = layout(nullptr, Rs, Qs);
    if (!workspace || workspace_bytes < need.bytes)
        return fail(CS_ERR_WORKSPACE, "workspace of %zu bytes < %zu required", workspace_bytes,
                    need.bytes);
    const Layout L = layout(workspace, Rs, Qs);
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const SpecView R = view(Rs), Q = view(Qs);
    const float mzmax = Rs.max_abs_mz > Qs.max_abs_mz ? Rs.max_abs_mz : Qs.max_abs_mz;
    const float tolc = coarse_tolerance(tol, shift, mzmax);
    const float shift32 = (float)shift;
    const int maxbin = Qs.max_peaks;
    auto blocks = [](int64_t n, int per) {
        int64_t b = (n + per - 1) / per;
        return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
    };

    const int stages = params->stages ? params->stages
                                      : (CS_STAGE_PREPARE | CS_STAGE_LIGHT | CS_STAGE_HEAVY);
    if (stages & ~(CS_STAGE_PREPARE | CS_STAGE_LIGHT | CS_STAGE_HEAVY))
        return fail(CS_ERR_INVALID, "unknown stage bits 0x%x", stages);
    if (stages & CS_STAGE_PREPARE) {
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(L.nbins, 256)), dim3(256), 0, hs, L.hist, L.nbins,
                       0);
    CS_LAUNCH_CHECK("k_fill_i32");
    hipLaunchKernelGGL(k_len_hist, dim3(blocks(nQ, 256)), dim3(256), 0, hs, Q.off, nQ, maxbin,
                       L.hist);
    CS_LAUNCH_CHECK("k_len_hist");
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(1024), 0, hs, L.hist, L.nbins, (int64_t)1,
                       (int64_t *)nullptr, L.cursor);
    CS_LAUNCH_CHECK("k_scan_block");
    hipLaunchKernelGGL(k_len_scatter, dim3(blocks(nQ, 256)), dim3(256), 0, hs, Q.off, nQ, maxbin,
                       L.cursor, L.order);
    CS_LAUNCH_CHECK("k_len_scatter");
    hipLaunchKernelGGL(k_tile_len, dim3(blocks(L.ntiles, 256)), dim3(256), 0, hs, Q.off, L.order,
                       nQ, L.ntiles, maxbin, L.tile_len);
    CS_LAUNCH_CHECK("k_tile_len");
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(1024), 0, hs, L.tile_len, L.ntiles,
                       (int64_t)kWave, L.tile_base, (int32_t *)nullptr);
    CS_LAUNCH_CHECK("k_scan_block");
    hipLaunchKernelGGL(k_tile_build, dim3(blocks(L.ntiles, 1)), dim3(256), 0, hs, Q.mz, Q.off,
                       L.order, nQ, L.tile_base, L.tile_len, L.ntiles, shift32, L.tiles);
    CS_LAUNCH_CHECK("k_tile_build");
    }

    const int64_t nrb = (nR + kRefsPerBlock - 1) / kRefsPerBlock;
    if (nrb > INT_MAX || L.ntiles > 65535)
        return fail(CS_ERR_UNSUPPORTED, "grid too large (%lld ref blocks, %lld query tiles)",
                    (long long)nrb, (long long)L.ntiles);
    const dim3 grid((unsigned)nrb, (unsigned)L.ntiles);
    const int64_t ml = params->match_limit;
    if (stages & CS_STAGE_LIGHT) {
    // the heavy-pair queue of this call starts empty
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, hs, L.counters, (int64_t)4);
    CS_LAUNCH_CHECK("k_fill_u64");
    switch (K) {
        case 1: st = launch_light<1>(grid, hs, R, nR, Q, nQ, L, tolc, tol, shift, ml, score, matches, overflow, ld); break;
        case 4: st = launch_light<4>(grid, hs, R, nR, Q, nQ, L, tolc, tol, shift, ml, score, matches, overflow, ld); break;
        case 16: st = launch_light<16>(grid, hs, R, nR, Q, nQ, L, tolc, tol, shift, ml, score, matches, overflow, ld); break;
        default: st = launch_light<8>(grid, hs, R, nR, Q, nQ, L, tolc, tol, shift, ml, score, matches, overflow, ld); break;
    }
    if (st != CS_OK) return st;
    }
    if (!(stages & CS_STAGE_HEAVY)) return CS_OK;

    HeavyArgs a{};
    a.R = R;
    a.Q = Q;
    a.nR = nR;
    a.nQ = nQ;
    a.tolc = tolc;
    a.shift32 = shift32;
    a.tol = tol;
    a.shift = shift;
    a.match_limit = ml;
    a.round_cap = T;
    int cap = 1;
    while (cap < 2 * T + kWave) cap <<= 1;
    a.buf_cap = cap;
    a.r_words = (Rs.max_peaks + 31) / 32 + 1;
    a.q_words = (Qs.max_peaks + 31) / 32 + 1;
    a.score = score;
    a.matches = matches;
    a.overflow = overflow;
    a.ld = ld;
    a.heavy_queue = L.heavy_queue;
    a.heavy_cap = L.heavy_cap;
    a.counters = L.counters;
    const size_t lds = (size_t)cap * (sizeof(double) + sizeof(uint32_t)) +
                       sizeof(uint32_t) * (size_t)(a.r_words + a.q_words);
    if (lds > 65536)
        return fail(CS_ERR_UNSUPPORTED,
                    "round_capacity %d with max_peaks %d/%d needs %zu bytes of LDS (> 64 KiB)", T,
                    Rs.max_peaks, Qs.max_peaks, lds);
    hipLaunchKernelGGL(k_heavy, dim3(kHeavyGrid), dim3(kWave), lds, hs, a);
    CS_LAUNCH_CHECK("k_heavy");
    return CS_OK;
}
