// cosine_greedy.hip -- all-pairs CosineGreedy on gfx950.
//
// Replaces the Numba-CUDA kernel of cudams/similarity/spectrum_similarity_
// functions.py:39-262 and the per-tile host loop of CudaCosineGreedy.py:149-185.
//
// Pipeline (all asynchronous on the caller's stream, no host sync):
//   1. k_fill_u64        zero the counters (a kernel, not a memset node, so the
//                        sequence can sit in a captured graph)
//   2. k_tile_meta, k_tile_index, k_tile_sort, k_tile_values
//                        per 64-query tile: every peak's m/z in one sorted
//                        list with its (query lane, peak) and a bucket
//                        directory over the tile's m/z range
//   3. k_light<K, F32>   lane l of a wave = pair (reference, query l of the
//                        tile): each reference peak finds its window in the
//                        tile list (directory + binary search), exact
//                        candidates go to the wave's LDS pool; pairs with at
//                        most K candidates settle in registers, up to 64 by
//                        the whole wave, more are queued
//   4. k_heavy           queued pairs: the small ones (at most 256 candidates
//                        and peaks) one per wave, then one workgroup per
//                        pair, largest size class first: enumeration into
//                        LDS, bitonic sort by the (v desc, i desc, j desc)
//                        key, one-wave greedy
//   5. k_giant           one wave per pair above the heavy buffer, in rounds
//                        of at most T candidates (exact in bounded LDS)
#include <limits.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "cs_internal.h"
#include "cg_lane.h"
#include "cg_tile.h"

namespace cs {
namespace {

constexpr int kWave = 64;
constexpr int kLightWaves = 16;
constexpr int kLightThreads = kWave * kLightWaves;
constexpr int kMaxLightRefs = 2048;   // largest refs_per_block
constexpr int kDefaultK = 8;
constexpr int kDefaultRound = 256;
constexpr int kHeavyGrid = 4096;
constexpr int64_t kHeavyQueueCap = int64_t(1) << 20;
constexpr int kMaxPeaks = 65535;
constexpr int kIndexStage = 12288;    // tiles up to this many peaks are indexed in LDS

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
// Prep kernels: query tiles and their peak index
// ---------------------------------------------------------------------------

__global__ void k_fill_u64(unsigned long long *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// The query set of one job as the tile kernels see it.  Global tile g of a
// batch belongs to job tile_job[g] and is that job's tile g - tile_base; its
// sorted list starts at pos_base + (the job's peaks before the tile) + (its
// local tile index), one +inf pad per tile.  A single call passes one job and
// no tile_job array.
struct PrepJob {
    const float *mz;
    const int64_t *off;
    const double *inp;
    int64_t nq;
    float shift32;
    int64_t tile_base;
    int64_t pos_base;
};

__device__ __forceinline__ const PrepJob &prep_job(const PrepJob &one, const PrepJob *jobs,
                                                  const int32_t *tile_job, int64_t g,
                                                  int64_t &t) {
    if (!tile_job) {
        t = g;
        return one;
    }
    const PrepJob &j = jobs[tile_job[g]];
    t = g - j.tile_base;
    return j;
}

// Per tile: peak count, position of its sorted list, value range.  One wave
// per tile, lane l reading query l's first and last peak; the range is a
// wave min / max (the per-query lists are sorted ascending).
__global__ void __launch_bounds__(256)
k_tile_meta(PrepJob one, const PrepJob *jobs, const int32_t *tile_job, int64_t ntiles,
            int32_t *tile_np, int64_t *tile_pbase, float *tile_lo, float *tile_iw) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t g = (int64_t)blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave); g < ntiles;
         g += waves) {
        int64_t t;
        const PrepJob &J = prep_job(one, jobs, tile_job, g, t);
        const float *mz = J.mz;
        const int64_t *off = J.off;
        const int64_t q0 = t * kTileQueries;
        const int64_t q1 = q0 + kTileQueries < J.nq ? q0 + kTileQueries : J.nq;
        const int64_t q = q0 + lane;
        float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
        if (q < q1) {
            const int64_t a = off[q], b = off[q + 1];
            if (b > a) {
                lo = mz[a] + J.shift32;
                hi = mz[b - 1] + J.shift32;
            }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const float ol = __shfl_xor(lo, o), oh = __shfl_xor(hi, o);
            lo = ol < lo ? ol : lo;
            hi = oh > hi ? oh : hi;
        }
        if (lane == 0) {
            if (!(lo <= hi)) lo = hi = 0.0f;
            tile_np[g] = (int32_t)(off[q1] - off[q0]);
            tile_pbase[g] = J.pos_base + (off[q0] - off[0]) + t;
            tile_lo[g] = lo;
            tile_iw[g] = tile_inv_width(lo, hi);
        }
    }
}

// Merge the tile's sorted per-query lists into one list sorted by
// (value, lane, peak): each element's rank is its own index plus, for every
// other list, the count of elements ordered before it (binary search).  The
// tile's values are staged in LDS when they fit.
constexpr int kSortStage = 8192;

template <class X>
__device__ void tile_rank_scatter(X xv, const int64_t *loff, int nl, int np, int64_t pb,
                                  float *smz, uint32_t *swho) {
    const int64_t g0 = loff[0];
    for (int e = threadIdx.x; e < np; e += blockDim.x) {
        const int64_t g = g0 + e;
        int ql = 0;
        while (ql + 1 < nl && loff[ql + 1] <= g) ++ql;
        const int j = (int)(g - loff[ql]);
        const float x = xv(e);
        int64_t rank = j;
        for (int l = 0; l < nl; ++l) {
            if (l == ql) continue;
            int a = (int)(loff[l] - g0), b = (int)(loff[l + 1] - g0);
            const int base = a;
            if (l < ql) {  // count y <= x
                while (a < b) {
                    const int m = (a + b) >> 1;
                    if (xv(m) <= x) a = m + 1;
                    else b = m;
                }
            } else {       // count y < x
                while (a < b) {
                    const int m = (a + b) >> 1;
                    if (xv(m) < x) a = m + 1;
                    else b = m;
                }
            }
            rank += a - base;
        }
        smz[pb + rank] = x;
        swho[pb + rank] = tile_who(ql, j);
    }
}

// Tiles above kIndexStage peaks: the list by rank merge, then its directory,
// dir[b] = first list position whose bucket is >= b for b in [0, kTileBuckets]
// (read back from the list this workgroup has just written).
__global__ void __launch_bounds__(256)
k_tile_sort(PrepJob one, const PrepJob *jobs, const int32_t *tile_job, int64_t ntiles,
            const int32_t *tile_np, const int64_t *tile_pbase, const float *tile_lo,
            const float *tile_iw, float *smz, uint32_t *swho, int32_t *dir) {
    __shared__ int64_t loff[kTileQueries + 1];
    __shared__ float xs[kSortStage];
    for (int64_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
        int64_t t;
        const PrepJob &J = prep_job(one, jobs, tile_job, g, t);
        const float shift32 = J.shift32;
        const int64_t q0 = t * kTileQueries;
        const int nl = (int)(q0 + kTileQueries < J.nq ? kTileQueries : J.nq - q0);
        if ((int)threadIdx.x <= nl) loff[threadIdx.x] = J.off[q0 + threadIdx.x];
        __syncthreads();
        const int np = tile_np[g];
        if (np <= kIndexStage) {  // uniform: k_tile_index handles it
            __syncthreads();
            continue;
        }
        const int64_t pb = tile_pbase[g];
        const float *src = J.mz + loff[0];
        if (np <= kSortStage) {
            for (int e = threadIdx.x; e < np; e += blockDim.x) xs[e] = src[e] + shift32;
            __syncthreads();
            tile_rank_scatter([&](int e) { return xs[e]; }, loff, nl, np, pb, smz, swho);
        } else {
            tile_rank_scatter([&](int e) { return src[e] + shift32; }, loff, nl, np, pb, smz,
                              swho);
        }
        if (threadIdx.x == 0) {
            smz[pb + np] = __builtin_huge_valf();
            swho[pb + np] = 0u;
        }
        __syncthreads();
        const float *s = smz + pb;
        const float lo = tile_lo[g], iw = tile_iw[g];
        int32_t *d = dir + g * (kTileBuckets + 1);
        for (int p = threadIdx.x; p <= np; p += blockDim.x) {
            const int bp = p < np ? tile_bucket(s[p], lo, iw) : kTileBuckets;
            const int bq = p > 0 ? tile_bucket(s[p - 1], lo, iw) : -1;
            for (int b = bq + 1; b <= bp; ++b) d[b] = p;
        }
        __syncthreads();
    }
}

// Tile index in LDS for tiles of at most kIndexStage peaks: counting sort by
// bucket (the exclusive prefix of the bucket histogram is the directory),
// then each element's final place inside its bucket is its rank by
// (value, lane, peak) among the bucket's elements.  Larger tiles are marked
// in `big` and indexed by k_tile_sort.
constexpr int kIndexThreads = 1024;
static_assert(kIndexStage % kIndexThreads == 0, "k_tile_index holds a tile's peaks in registers");

__global__ void __launch_bounds__(kIndexThreads)
k_tile_index(PrepJob one, const PrepJob *jobs, const int32_t *tile_job, int64_t ntiles,
             const int32_t *tile_np, const int64_t *tile_pbase, const float *tile_lo,
             const float *tile_iw, float *smz, uint32_t *swho, int32_t *dir) {
    extern __shared__ __align__(16) unsigned char index_lds[];
    __shared__ int64_t loff[kTileQueries + 1];
    __shared__ int part[kIndexThreads];
    int *start = reinterpret_cast<int *>(index_lds);     // kTileBuckets + 1
    int *cursor = start + kTileBuckets + 1;               // kTileBuckets
    float *bx = reinterpret_cast<float *>(cursor + kTileBuckets);
    uint32_t *bw = reinterpret_cast<uint32_t *>(bx + kIndexStage);
    const int tid = threadIdx.x;
    for (int64_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
        const int np = tile_np[g];
        if (np > kIndexStage) continue;  // uniform: left to the rank-merge path
        int64_t t;
        const PrepJob &J = prep_job(one, jobs, tile_job, g, t);
        const float *mz = J.mz;
        const float shift32 = J.shift32;
        const int64_t q0 = t * kTileQueries;
        const int nl = (int)(q0 + kTileQueries < J.nq ? kTileQueries : J.nq - q0);
        if (tid <= nl) loff[tid] = J.off[q0 + tid];
        for (int b = tid; b <= kTileBuckets; b += kIndexThreads) start[b] = 0;
        __syncthreads();
        const float lo = tile_lo[g], iw = tile_iw[g];
        const int64_t g0 = loff[0];
        // the thread's peaks (e = tid + u * kIndexThreads), loaded once for
        // the histogram and the scatter: one memory latency, not one per peak
        constexpr int kPer = kIndexStage / kIndexThreads;
        float xr[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * kIndexThreads;
            xr[u] = e < np ? mz[g0 + e] + shift32 : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            if (tid + u * kIndexThreads < np) atomicAdd(&start[tile_bucket(xr[u], lo, iw)], 1);
        __syncthreads();
        // exclusive scan of start[0..kTileBuckets) (8 buckets per thread)
        int loc[kTileBuckets / kIndexThreads];
        int sum = 0;
#pragma unroll
        for (int u = 0; u < kTileBuckets / kIndexThreads; ++u) {
            loc[u] = start[tid * (kTileBuckets / kIndexThreads) + u];
            sum += loc[u];
        }
        // inclusive scan of the thread sums: within each wave by shuffles,
        // then over the 16 wave totals
        const int lane = tid & (kWave - 1), wv = tid / kWave;
        int incl = sum;
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == kWave - 1) part[wv] = incl;
        __syncthreads();
        if (tid < kWave) {
            int w = tid < kIndexThreads / kWave ? part[tid] : 0;
            for (int o = 1; o < kIndexThreads / kWave; o <<= 1) {
                const int y = __shfl_up(w, o);
                if (tid >= o) w += y;
            }
            if (tid < kIndexThreads / kWave) part[kIndexThreads / kWave + tid] = w;
        }
        __syncthreads();
        int run = incl - sum + (wv ? part[kIndexThreads / kWave + wv - 1] : 0);
#pragma unroll
        for (int u = 0; u < kTileBuckets / kIndexThreads; ++u) {
            const int b = tid * (kTileBuckets / kIndexThreads) + u;
            start[b] = run;
            cursor[b] = run;
            run += loc[u];
        }
        if (tid == 0) start[kTileBuckets] = np;
        __syncthreads();
        int32_t *d = dir + g * (kTileBuckets + 1);
        for (int b = tid; b <= kTileBuckets; b += kIndexThreads) d[b] = start[b];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {  // scatter into buckets
            const int e = tid + u * kIndexThreads;
            if (e >= np) break;
            const int64_t gp = g0 + e;
            int ql = 0;  // last list whose offset is <= gp (binary search)
            for (int step = kTileQueries / 2; step > 0; step >>= 1)
                if (ql + step < nl && loff[ql + step] <= gp) ql += step;
            const float x = xr[u];
            const int pos = atomicAdd(&cursor[tile_bucket(x, lo, iw)], 1);
            bx[pos] = x;
            bw[pos] = tile_who(ql, (int)(gp - loff[ql]));
        }
        __syncthreads();
        const int64_t pb = tile_pbase[g];
        for (int e = tid; e < np; e += kIndexThreads) {  // rank inside the bucket
            const float x = bx[e];
            const uint32_t w = bw[e];
            const int b = tile_bucket(x, lo, iw);
            // rank = bucket start + entries of the bucket ordered before
            // (x, w); eight loads in flight per round (a peak cluster puts
            // about 100 entries in one bucket)
            int rank = start[b];
            int k = start[b];
            const int kend = start[b + 1];
            for (; k + 8 <= kend; k += 8) {
                float xs[8];
                uint32_t ws[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    xs[u] = bx[k + u];
                    ws[u] = bw[k + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) rank += (xs[u] < x) | ((xs[u] == x) & (ws[u] < w));
            }
            for (; k < kend; ++k) rank += (bx[k] < x) | ((bx[k] == x) & (bw[k] < w));
            smz[pb + rank] = x;
            swho[pb + rank] = w;
        }
        if (tid == 0) {
            smz[pb + np] = __builtin_huge_valf();
            swho[pb + np] = 0u;
        }
        __syncthreads();
    }
}

// sint[p] = float32 of the int_pow of list position p, and whether every
// value of the tile survives that rounding (then the lane kernel forms
// candidate values from these copies exactly).
//
// With tile_order, the grid's last workgroup instead writes the order in
// which a scheduled lane-kernel launch takes the tiles: most peaks first (a
// tile's lane-kernel time follows its peak count), ties by tile index, so
// the longest work items start first and the last round is made of the
// shortest.  It ranks (peaks, ~index) keys in LDS.  Scheduled
// launches have at most kOrderTiles tiles; more tiles leave the order
// unwritten (it is not read).
constexpr int kOrderTiles = 2048;
__device__ void order_tiles(const int32_t *tile_np, int64_t ntiles, uint32_t *order) {
    __shared__ unsigned long long key[kOrderTiles];
    if (ntiles > kOrderTiles) return;
    const int tid = threadIdx.x, nt = blockDim.x, n = (int)ntiles;
    for (int t = tid; t < n; t += nt)
        key[t] = ((unsigned long long)(uint32_t)tile_np[t] << 32) |
                 (unsigned long long)(0xffffffffu - (uint32_t)t);
    __syncthreads();
    // a tile's place is the number of greater keys (the keys are distinct);
    // every lane reads the same key at once, so the loop is LDS broadcasts
    for (int t = tid; t < n; t += nt) {
        const unsigned long long k = key[t];
        int rank = 0;
        for (int u = 0; u < n; ++u) rank += key[u] > k ? 1 : 0;
        order[rank] = (uint32_t)t;
    }
}

__global__ void __launch_bounds__(1024)
k_tile_values(PrepJob one, const PrepJob *jobs, const int32_t *tile_job, int64_t ntiles,
              const int32_t *tile_np, const int64_t *tile_pbase, const uint32_t *swho,
              float *sint, int32_t *tile_flags, uint32_t *tile_order) {
    __shared__ int64_t loff[kTileQueries + 1];
    const int tid = threadIdx.x;
    const int64_t nblk = tile_order ? gridDim.x - 1 : gridDim.x;
    if (blockIdx.x == nblk) {  // the extra workgroup of a tile_order launch
        order_tiles(tile_np, ntiles, tile_order);
        return;
    }
    for (int64_t g = blockIdx.x; g < ntiles; g += nblk) {
        int64_t t;
        const PrepJob &J = prep_job(one, jobs, tile_job, g, t);
        const double *inp = J.inp;
        const int64_t q0 = t * kTileQueries;
        const int nl = (int)(q0 + kTileQueries < J.nq ? kTileQueries : J.nq - q0);
        if (tid <= nl) loff[tid] = J.off[q0 + tid];
        __syncthreads();
        const int np = tile_np[g];
        const int64_t pb = tile_pbase[g];
        int exact = 1;
        constexpr int kBatch = 8;  // loads in flight per thread
        for (int p0 = tid; p0 < np; p0 += kBatch * (int)blockDim.x) {
            uint32_t w[kBatch];
            double x[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int p = p0 + u * (int)blockDim.x;
                w[u] = p < np ? swho[pb + p] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int p = p0 + u * (int)blockDim.x;
                x[u] = p < np ? inp[loff[w[u] >> 16] + (w[u] & 0xffffu)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int p = p0 + u * (int)blockDim.x;
                if (p < np) {
                    const float f = (float)x[u];
                    sint[pb + p] = f;
                    exact &= (double)f == x[u] ? 1 : 0;
                }
            }
        }
        if (tid == 0) sint[pb + np] = 0.0f;  // the pad position
        exact = __syncthreads_and(exact);
        if (tid == 0) tile_flags[g] = exact;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Light kernel: one wave per (reference, tile), lanes over reference peaks
// ---------------------------------------------------------------------------

struct LightArgs {
    SpecView R, Q;
    int64_t nR, nQ;
    const int32_t *tile_np;
    const int64_t *tile_pbase;
    const float *tile_lo, *tile_iw;
    const float *smz;
    const uint32_t *swho;
    const float *sint;         // query int_pow in sorted-list order, as float32
    const int32_t *tile_flags; // bit 0: every sint value of the tile is exact
    const int32_t *dir;
    int lds_cap;               // sorted peaks a tile may hold in LDS
    float tolc;
    double tol, shift;
    int exact_from_tile;       // query m/z in the sorted list are the exact values
    int no_mz_pow;             // mz_power == 0 on both sides (mz_pow arrays NULL)
    int64_t match_limit;
    float *score;
    int32_t *matches;
    uint8_t *overflow;
    int64_t ld;
    unsigned long long *heavy_queue;
    int64_t heavy_cap;
    unsigned long long *counters;
    int64_t n_ref_blocks;      // grid = n_ref_blocks x tiles, reference block fastest
    int refs_per_block;        // references one workgroup walks (<= kMaxLightRefs)
    int job;                   // this job's index in a batch
    int batched;               // heavy-queue entries carry the job index
};

// A queued pair.  One call: r << 32 | q.  A batch: job << 52 | r << 26 | q
// (jobs < 4096, references and queries < 2^26 per job).
__device__ __forceinline__ unsigned long long pair_entry(int batched, int job, int64_t r,
                                                        int64_t q) {
    if (batched)
        return ((unsigned long long)job << 52) | ((unsigned long long)r << 26) |
               (unsigned long long)q;
    return ((unsigned long long)r << 32) | (unsigned int)q;
}

struct PairRef {
    int job;
    int64_t r, q;
};

__device__ __forceinline__ PairRef pair_of(int batched, unsigned long long e) {
    if (batched)
        return PairRef{(int)(e >> 52), (int64_t)((e >> 26) & 0x3ffffffull),
                       (int64_t)(e & 0x3ffffffull)};
    return PairRef{0, (int64_t)(e >> 32), (int64_t)(e & 0xffffffffull)};
}

// A value every lane of the wave holds equally, moved to a scalar register
// so that branches and loop bounds on it are scalar.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Value of lane f (wave-uniform f) without an LDS round trip.
__device__ __forceinline__ int lane_int(int x, int f) { return __builtin_amdgcn_readlane(x, f); }
__device__ __forceinline__ double lane_double(double x, int f) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, f);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), f);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Orders LDS accesses of the lanes of one wave.  LDS instructions of a wave
// execute in issue order, so this only has to stop the compiler from moving
// accesses across it (no s_waitcnt on outstanding global stores).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Per-wave candidate pool of k_light: every exact candidate of the current
// reference against the tile, in emit order, with its query lane (pq).  The
// first K (light_candidates) candidates of each query are also indexed by
// their rank (pidx[rank][query], the rank is what the query's count atomic
// returns; one spare row takes every later rank, so the store needs no
// branch), so a lane gathers its pair's candidates with independent loads.  A
// query contributes at most kMidCap entries (more means the pair is settled
// by the heavy kernel anyway).
constexpr int kPool = 128;
// Pool arrays hold kPool entries and a spare (index kPool) that an emit with
// the pool full writes to instead of branching; pdrop likewise has a spare
// entry (index kTileQueries).  Nothing reads the spares.  kPool + 2 keeps the
// arrays after the float64 values 8-byte aligned.
constexpr int kPoolSlots = kPool + 2;
constexpr int kDropSlots = kTileQueries + 1;
constexpr int kMidCap = 64;     // candidates a wave settles for one pair in registers
// Heavy pairs settled by a whole workgroup go in four size classes, taken
// largest first so that the largest pairs do not form the kernel's tail:
// class 3 above 2048 candidates, 2 above 1024, 1 above 512, 0 the rest.
constexpr int kHeavyClasses = 4;
__device__ __forceinline__ int heavy_class(int n) {
    return n > 2048 ? 3 : (n > 1024 ? 2 : (n > 512 ? 1 : 0));
}
// Heavy pairs of at most kWaveKeys candidates between spectra of at most
// kWavePeaks peaks are settled by one wave each (eight pairs per heavy
// workgroup at once) instead of by a whole workgroup.
constexpr int kWavePeaks = 256;
constexpr int kWaveKeys = 256;
// Counters of one call (zeroed by k_fill_u64 before the lane kernel):
// [1] next workgroup pair, [2] giant pairs, [3] next giant pair, [5] [6]
// lane-kernel work items (scheduled / batch launches), [7] small pairs,
// [8] next small pairs, [9 + k] workgroup pairs of size class k.  The heavy
// queue's lower half holds four equal regions, one per size class, each
// filled from its bottom; its upper half holds the small pairs.
constexpr int kCounters = 16;
constexpr int kSmallCount = 7;
constexpr int kSmallNext = 8;
constexpr int kClassCount = 9;
// Slots of the block-relative reference offsets in LDS: kMaxLightRefs + 1
// rounded up so that the pools after them start 16-byte aligned (their
// float64 values are read and written as 8-byte LDS accesses).
constexpr int kRoffSlots = (kMaxLightRefs + 1 + 3) / 4 * 4;
constexpr size_t kPoolBase = (size_t)kTileQueries * (8 + 8) + (size_t)kRoffSlots * 4;
constexpr int kLightPeakBytes = 4 + 4 + 2 + 1;
// LDS of one gfx950 workgroup (and CU): the lane kernel is laid out for it at
// compile time, so every LDS array sits at a constant offset.
constexpr size_t kLdsBytes = 160 * 1024;

// LDS layout of the lane kernel for K = light_candidates.  The rank index
// has K rows and a spare.  The directory over the tile's sorted values has
// `buckets` buckets: twice the prepared directory's for K <= 8, paid for by
// the rank rows those variants do not need, so the probe's binary search
// starts from half as many entries.  A tile of at most `cap` peaks is staged
// in LDS; larger tiles are probed in HBM through the prepared directory.
template <int K>
struct LightLayout {
    static constexpr int ranks = K;
    static constexpr int buckets = K <= 8 ? 2 * kTileBuckets : kTileBuckets;
    static constexpr size_t wave_bytes =
        ((size_t)kPoolSlots * (8 + 4 + 1) + (size_t)kTileQueries * (4 + ranks + 1) +
         kDropSlots + 16 + kWave * 4 + 15) / 16 * 16;
    static constexpr int dir_pad = (buckets + 1 + 7) / 8 * 8;
    // bytes of dynamic LDS for a tile capacity `c`
    static constexpr size_t bytes(int c) {
        return (size_t)kTileQueries * (8 + 8)          // qoff, qnorm
               + (size_t)kRoffSlots * 4                 // reference offsets (block-relative)
               + (size_t)kLightWaves * wave_bytes       // pools
               + (size_t)(c + 1) * 4                    // sorted values (f32)
               + (size_t)(c + 1) * 4                    // sorted int_pow (f32)
               + (size_t)dir_pad * 2                    // directory (u16)
               + (size_t)(c + 1) * 2                    // peak index (u16)
               + (size_t)(c + 1 + 15) / 16 * 16;        // query lane (u8)
    }
    static constexpr int cap = (int)((kLdsBytes - bytes(0) - 64) / kLightPeakBytes) - 1;
    static_assert(kPoolBase % 16 == 0 && wave_bytes % 16 == 0, "pool values are 8-byte aligned");
    static_assert(bytes(cap) <= kLdsBytes, "lane kernel LDS fits a workgroup");
    static_assert(cap < 65535, "directory entries are 16-bit");
};

// Bucket of x among `nb` equal buckets over the tile's values: nb =
// kTileBuckets << s uses inv_w << s (exact), so it refines the prepared
// directory's buckets and stays monotone in x.
constexpr int kMaxLightBuckets = 32768;
constexpr int kDirBuildRefs = 512;  // fewer references per workgroup copy the prepared directory
__device__ __forceinline__ int light_bucket(float x, float lo, float inv_w_nb, int nb) {
    const float f = (x - lo) * inv_w_nb;
    if (!(f > 0.0f)) return 0;
    if (f >= (float)(nb - 1)) return nb - 1;
    return (int)f;
}

// Descending bitonic sort of one key per lane across the wave (64 lanes),
// partner values from xor_lane (cs_internal.h).

// One compare-exchange stage of the bitonic network (block size Kb, partner
// distance J).
template <int Kb, int J>
__device__ __forceinline__ void sort_stage(double &v, uint32_t &ij, int lane) {
    const double ov = xor_lane_double<J>(v);
    const uint32_t oij = (uint32_t)xor_lane<J>((int)ij);
    const bool lower = (lane & J) == 0;
    const bool desc = (lane & Kb) == 0;
    // the lower index of a descending block keeps the greater key
    const bool keep_greater = lower == desc;
    const bool other_greater = key_greater(ov, oij, v, ij);
    const bool mine_greater = key_greater(v, ij, ov, oij);
    const bool take = keep_greater ? other_greater : mine_greater;
    v = take ? ov : v;
    ij = take ? oij : ij;
}

template <int Kb, int J>
__device__ __forceinline__ void sort_merge(double &v, uint32_t &ij, int lane) {
    sort_stage<Kb, J>(v, ij, lane);
    if constexpr (J > 1) sort_merge<Kb, J / 2>(v, ij, lane);
}

// Bitonic sort of one (value, ij) per lane across the wave, descending by key,
// for keys in lanes [0, n) and -inf keys above (n wave-uniform).  Lanes
// [0, 16) sorted by the 16-lane merges hold all n <= 16 keys in descending
// order, so the wider merges run only for larger n.
__device__ __forceinline__ void wave_sort_desc(double &v, uint32_t &ij, int n) {
    const int lane = threadIdx.x & (kWave - 1);
    sort_merge<2, 1>(v, ij, lane);
    sort_merge<4, 2>(v, ij, lane);
    sort_merge<8, 4>(v, ij, lane);
    sort_merge<16, 8>(v, ij, lane);
    if (n > 16) {
        sort_merge<32, 16>(v, ij, lane);
        if (n > 32) sort_merge<64, 32>(v, ij, lane);
    }
}

// s_waitcnt operand: vmcnt(0), expcnt and lgkmcnt not waited on (gfx9 layout:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
constexpr int kWaitVmcnt0 = 0x0F70;

// The first two 64-peak chunks of one reference, held in registers from one
// step before they are probed (lane l holds peaks l and 64 + l).
struct RefChunk {
    int rr;        // reference within the workgroup's block (>= nrefs: none)
    int n;
    int64_t off;
    double norm;
    float mz0, mz1;
    double in0, in1;
};

// LDS layout of the lane kernel's per-wave candidate pool (see
// LightLayout): values, (i, j), per-query counts, pool count, compaction
// slots, query of each slot.
struct WavePool {
    double *pv;
    uint32_t *pij;
    int *cnt;
    int *pool_n;
    int *midx;
    unsigned char *pq;
};

// The pool of the wave whose pool starts `pool_off` bytes into the LDS.
__device__ __forceinline__ WavePool wave_pool(unsigned char *lds, int pool_off) {
    WavePool w;
    w.pv = reinterpret_cast<double *>(lds + pool_off);
    w.pij = reinterpret_cast<uint32_t *>(w.pv + kPoolSlots);
    w.cnt = reinterpret_cast<int *>(w.pij + kPoolSlots);
    w.pool_n = w.cnt + kTileQueries;
    w.midx = w.pool_n + 4;
    w.pq = reinterpret_cast<unsigned char *>(w.midx + kWave);
    return w;
}

struct MidResult {
    double sc;  // greedy score sum (not yet normalised)
    int m;      // matches
};

// One pair with K < n <= kMidCap candidates in the wave's pool (slots
// 0 .. lim - 1, owner lane sl), settled by the whole wave: the pair's slots
// are compacted, their (value, i, j) sorted descending across the lanes, and
// the greedy pass accepts the first live candidate, then drops every one that
// shares its i or j.  The same result on every lane.
__device__ __attribute__((noinline)) MidResult mid_settle(int pool_off, int lim, int sl) {
    extern __shared__ __align__(16) unsigned char light_lds[];
    const WavePool P = wave_pool(light_lds, pool_off);
    const int lane = threadIdx.x & (kWave - 1);
    int have = 0;
    for (int base = 0; base < lim; base += kWave) {  // compact the pair's entries
        const int e = base + lane;
        const bool mine_e = e < lim && P.pq[e] == sl;
        const unsigned long long mk = __ballot(mine_e);
        // at most kMidCap entries belong to sl, so every slot is < 64
        if (mine_e) P.midx[have + __popcll(mk & ((1ull << lane) - 1ull))] = e;
        have += __popcll(mk);
    }
    wave_sync();
    double v = -__builtin_huge_val();
    uint32_t ij = 0u;
    if (lane < have) {
        const int e = P.midx[lane];
        v = P.pv[e];
        ij = P.pij[e];
    }
    wave_sync();
    wave_sort_desc(v, ij, have);
    const int vi = (int)(ij >> 16), vj = (int)(ij & 0xffffu);
    bool alive = lane < have;
    MidResult r{0.0, 0};  // the same on every lane (read from lane f)
    unsigned long long mk = __ballot(alive);
    while (mk) {
        const int f = __builtin_ctzll(mk);
        const int fi = lane_int(vi, f), fj = lane_int(vj, f);
        r.sc += lane_double(v, f);
        ++r.m;
        alive = alive & (lane != f) & (vi != fi) & (vj != fj);
        mk = __ballot(alive);
    }
    return r;
}

// One workgroup = one 64-query tile x refs_per_block references.  The tile's peak
// index sits in LDS (shared by 16 waves); wave w settles references taken
// from a workgroup counter one at a time: lanes over the reference's peaks
// probe the index, exact candidates go to the wave's pool, then lane l
// settles the pair (reference, query l).  Each reference's first 128 peaks
// and int_pow are loaded a whole step before it is probed (two register
// buffers used in turn).
//
// F32: every m/z is its float32 value on both sides, there is no shift and
// mz_power is 0.  Then the exact window test needs only the float32 values
// at hand, and a candidate's value ((1 * int_pow_r) * 1) * int_pow_q comes
// from registers and LDS when the tile's int_pow values are float32 numbers,
// so a step waits on no global load.  The general variant reads the float64
// arrays it needs.
template <int K, bool F32>
__device__ void light_body(const LightArgs &a, int64_t rb, int64_t t) {
    extern __shared__ __align__(16) unsigned char light_lds[];
    __shared__ int next_ref;
    using LL = LightLayout<K>;
    constexpr int cap = LL::cap;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // LDS layout (see LightLayout)
    int64_t *qoff_l = reinterpret_cast<int64_t *>(light_lds);
    double *qnorm_l = reinterpret_cast<double *>(qoff_l + kTileQueries);
    int32_t *roff_l = reinterpret_cast<int32_t *>(qnorm_l + kTileQueries);
    unsigned char *pools = reinterpret_cast<unsigned char *>(roff_l + kRoffSlots);
    unsigned char *wbase = pools + wave * LL::wave_bytes;
    const int pool_off = (int)(wbase - light_lds);
    double *pv = reinterpret_cast<double *>(wbase);
    uint32_t *pij = reinterpret_cast<uint32_t *>(pv + kPoolSlots);
    int *cnt = reinterpret_cast<int *>(pij + kPoolSlots);
    int *pool_n = cnt + kTileQueries;
    int *midx = pool_n + 4;
    unsigned char *pq = reinterpret_cast<unsigned char *>(midx + kWave);
    unsigned char *pdrop = pq + kPoolSlots;  // query lost a candidate to a full pool
    unsigned char *pidx = pdrop + kDropSlots;  // [K + 1][kTileQueries] pool slots
    // The staged tile is laid out for its own peak count (np + 1 with the +inf
    // pad), and the directory takes the rest: at least LL::buckets buckets
    // (what the capacity was computed with), doubled while it fits, so a
    // typical tile gets a finer one.  The probe computes these addresses from
    // scalar bases, as it would from constant ones.
    const int np = a.tile_np[t];
    const int n1 = np + 1;
    float *smz_l = reinterpret_cast<float *>(pools + kLightWaves * LL::wave_bytes);
    float *sint_l = smz_l + n1;
    uint16_t *wj_l = reinterpret_cast<uint16_t *>(sint_l + n1);
    unsigned char *wl_l = reinterpret_cast<unsigned char *>(wj_l + n1);
    uint16_t *dir_l = reinterpret_cast<uint16_t *>(wl_l + ((n1 + 1) & ~1));
    const int dir_room = (int)((int64_t)LL::bytes(cap) -
                               (int64_t)(reinterpret_cast<unsigned char *>(dir_l) - light_lds)) / 2;
    // A workgroup with few references copies the prepared directory of the
    // tile (kTileBuckets buckets, read with the tile) instead of building a
    // finer one: the build's divergent bucket fills cost about 4.4 us per
    // workgroup, as much as a few hundred references gain from the finer
    // directory.
    const int64_t r_first = rb * a.refs_per_block;
    const bool copy_dir =
        (r_first + a.refs_per_block < a.nR ? a.refs_per_block : a.nR - r_first) < kDirBuildRefs;
    int nb = copy_dir ? kTileBuckets : LL::buckets;  // uniform
    if (!copy_dir)
        while (2 * nb + 1 <= dir_room && 2 * nb <= kMaxLightBuckets) nb *= 2;
    const int64_t pb = a.tile_pbase[t];
    const float lo = a.tile_lo[t], iw = a.tile_iw[t];
    const float *gsmz = a.smz + pb;
    const uint32_t *gswho = a.swho + pb;
    const float *gsint = a.sint + pb;
    const int32_t *gdir = a.dir + t * (kTileBuckets + 1);
    const bool in_lds = np <= cap;
    // candidate values from float32 copies: exact when mz_power == 0 and the
    // tile's int_pow values are float32 numbers
    const bool tile_f32 = F32 && (a.tile_flags[t] & 1);
    const int64_t q0 = t * kTileQueries;
    const int64_t r0 = rb * a.refs_per_block;
    const int nrefs = (int)(r0 + a.refs_per_block < a.nR ? a.refs_per_block : a.nR - r0);
    const SpecView &R = a.R, &Q = a.Q;
    const int64_t rbase = R.off[r0];
    constexpr int kDirPer = (kTileBuckets + kLightThreads) / kLightThreads;
    int32_t dcopy[kDirPer];
    if (in_lds && copy_dir) {  // in flight with the tile's loads below
#pragma unroll
        for (int u = 0; u < kDirPer; ++u) {
            const int b = tid + u * kLightThreads;
            dcopy[u] = b <= kTileBuckets ? gdir[b] : 0;
        }
    }
    if (in_lds) {  // several independent loads in flight per thread
        for (int e = tid; e <= np; e += 4 * kLightThreads) {
            float x[4], y[4];
            uint32_t w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int f = e + u * kLightThreads;
                x[u] = f <= np ? gsmz[f] : 0.0f;
                y[u] = F32 && f <= np ? gsint[f] : 0.0f;
                w[u] = f <= np ? gswho[f] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int f = e + u * kLightThreads;
                if (f <= np) {
                    smz_l[f] = x[u];
                    if (F32) sint_l[f] = y[u];
                    wj_l[f] = (uint16_t)(w[u] & 0xffffu);
                    wl_l[f] = (unsigned char)(w[u] >> 16);
                }
            }
        }
    }
    if (in_lds && copy_dir) {
#pragma unroll
        for (int u = 0; u < kDirPer; ++u) {
            const int b = tid + u * kLightThreads;
            if (b <= kTileBuckets) dir_l[b] = (uint16_t)dcopy[u];
        }
    }
    if (tid < kTileQueries) {
        const int64_t q = q0 + tid < a.nQ ? q0 + tid : q0;
        qoff_l[tid] = Q.off[q];
        qnorm_l[tid] = Q.norm[q];
    }
    for (int x = tid; x <= nrefs; x += kLightThreads) roff_l[x] = (int32_t)(R.off[r0 + x] - rbase);
    cnt[lane] = 0;
    pdrop[lane] = 0;
    if (lane == 0) *pool_n = 0;
    if (tid == 0) next_ref = 0;
    __syncthreads();
    // The directory of the staged tile, built here from its sorted values:
    // dir[b] = the first position whose bucket is >= b (b in [0, buckets]).
    // Position p fills the buckets after its predecessor's, up to its own;
    // the +inf pad at np fills the rest.
    const float iwn = iw * (float)(nb / kTileBuckets);  // nb / kTileBuckets: a power of two
    if (in_lds && !copy_dir) {
        for (int p = tid; p <= np; p += kLightThreads) {
            const int bc = p < np ? light_bucket(smz_l[p], lo, iwn, nb) : nb;
            const int bp = p > 0 ? light_bucket(smz_l[p - 1], lo, iwn, nb) : -1;
            for (int b = bp + 1; b <= bc; ++b) dir_l[b] = (uint16_t)p;
        }
        __syncthreads();
    }

    const int64_t q = q0 + lane;
    const bool valid = q < a.nQ;
    const unsigned ulane = (unsigned)lane;

    auto grab = [&]() {
        int v = 0;
        if (lane == 0) v = atomicAdd(&next_ref, 1);
        return uniform(v);
    };
    auto fetch = [&](RefChunk &c, int rr) {
        c.rr = rr;
        c.n = 0;
        c.off = 0;
        c.norm = 1.0;
        if (rr < nrefs) {
            const int b0 = uniform(roff_l[rr]);
            c.off = rbase + b0;
            c.n = uniform(roff_l[rr + 1]) - b0;
            c.norm = R.norm[r0 + rr];
        }
        // uniform base + 32-bit lane offset: the scalar-base address form,
        // no per-lane 64-bit address arithmetic
        const float *rmz = R.mz + c.off;
        const double *rin = R.inp + c.off;
        const bool h0 = lane < c.n, h1 = lane + kWave < c.n;
        c.mz0 = h0 ? rmz[ulane] : 0.0f;
        c.mz1 = h1 ? rmz[kWave + ulane] : 0.0f;
        c.in0 = h0 ? rin[ulane] : 0.0;
        c.in1 = h1 ? rin[kWave + ulane] : 0.0;
    };

    // outputs of the previous reference, held in registers and stored only
    // after the next probe (a store still in flight would hold up the next
    // wait for the prefetched peaks)
    int64_t out_row = -1;
    float out_score = 0.0f;
    int out_matches = 0;
    int out_overflow = 0;
    auto flush = [&]() {
        if (out_row >= 0 && valid) {
            const int64_t o = out_row * a.ld + q0;  // uniform: the row's tile segment
            (a.score + o)[ulane] = out_score;
            (a.matches + o)[ulane] = out_matches;
            (a.overflow + o)[ulane] = (uint8_t)out_overflow;
        }
    };

    auto step = [&](const RefChunk &c) {
        const int rr = c.rr;
        const int64_t r = r0 + rr;
        const int64_t roff = c.off;
        const int nr = c.n;

        // 1. probe: lanes over the reference's peaks, exact candidates of the
        // queries in `filter` into the pool.  TF (tile float32 values) is a
        // compile-time branch around the whole probe, and the two register
        // chunks are probed outside the loop over further chunks, so the
        // common path has no global load and no wait in it.
        auto probe = [&](auto tf, auto from_regs, unsigned long long filter) {
            constexpr bool TF = decltype(tf)::value;
            constexpr bool REGS = decltype(from_regs)::value;
            auto chunk = [&](int ib, float ri, double rin) {
                const int i = ib + lane;
                if (i >= nr) return;
                // the walk covers the coarse float32 window [x0, x1]; the
                // float64 test of each entry narrows it to the exact one
                const float x0 = ri - a.tolc, x1 = ri + a.tolc;
                auto emit = [&](int ql, int j, float qv, float qin) {
                    if (!((filter >> ql) & 1ull)) return;
                    if constexpr (F32) {  // no shift on this path
                        if (!exact_window_noshift((double)ri, (double)qv, a.tol)) return;
                    } else {
                        const double mq = a.exact_from_tile ? (double)qv
                                                            : exact_mz(Q, qoff_l[ql] + j);
                        const double mr = R.mzx ? R.mzx[roff + i] : (double)ri;
                        if (!exact_window(mr, mq, a.tol, a.shift)) return;
                    }
                    // rank within the query and pool slot, both atomics in flight at once
                    const int rank = atomicAdd(&cnt[ql], 1);
                    const int pos = atomicAdd(pool_n, 1);
                    // The rare cases take no branch.  A pair past kMidCap
                    // candidates goes to the heavy kernel: its slot is marked
                    // empty.  With the pool full the entry goes to the spare
                    // slot and the query is probed again in a later pass
                    // (pdrop; a dropped pair reads nothing of this pass).
                    const bool room = pos < kPool, keep = rank < kMidCap;
                    const int slot = room ? pos : kPool;
                    if constexpr (F32 && TF)  // ((1 * int_pow_r) * 1) * int_pow_q
                        pv[slot] = rin * (double)qin;
                    // otherwise the value needs the query peak's arrays in
                    // HBM: filled for the whole pool after the probe
                    pij[slot] = pack_ij(i, j);
                    pq[slot] = keep ? (unsigned char)ql : (unsigned char)0xffu;
                    // ranks past the index go to a spare row nobody reads
                    pidx[(rank < LL::ranks ? rank : LL::ranks) * kTileQueries + ql] =
                        (unsigned char)slot;
                    pdrop[(room | !keep) ? kTileQueries : ql] = 1;
                };
                if (in_lds) {
                    // tile_probe on the LDS copy, with the skip done as a
                    // binary search inside the bucket: positions before
                    // dir[b] hold values below x0 and positions from dir[b+1]
                    // on hold values above it (bucket() is monotone), so the
                    // first value >= x0 sits in [dir[b], dir[b+1]].  About 3
                    // rounds per wave instead of about 8 for the linear walk.
                    const int b = light_bucket(x0, lo, iwn, nb);
                    int pl = (int)dir_l[b], ph = (int)dir_l[b + 1];
                    while (pl < ph) {
                        const int mid = (pl + ph) >> 1;
                        if (smz_l[mid] < x0)
                            pl = mid + 1;
                        else
                            ph = mid;
                    }
                    int p = pl;
                    float v = smz_l[p];
                    while (v <= x1) {
                        emit((int)wl_l[p], (int)wj_l[p], v, TF ? sint_l[p] : 0.0f);
                        v = smz_l[++p];
                    }
                } else {
                    tile_probe(x0, x1, lo, iw, [&](int p) { return gsmz[p]; },
                               [&](int b) { return (int)gdir[b]; },
                               [&](int p) {
                                   const uint32_t w = gswho[p];
                                   emit((int)(w >> 16), (int)(w & 0xffffu), gsmz[p],
                                        TF ? gsint[p] : 0.0f);
                               });
                }
            };
            int from = 0;
            if constexpr (REGS) {
                chunk(0, c.mz0, c.in0);
                if (nr > kWave) chunk(kWave, c.mz1, c.in1);
                from = 2 * kWave;
            }
            for (int ib = from; ib < nr; ib += kWave) {
                const int i = ib + lane;
                const float ri = i < nr ? R.mz[roff + i] : 0.0f;
                const double rin = i < nr ? R.inp[roff + i] : 0.0;
                chunk(ib, ri, rin);
            }
            if constexpr (!(F32 && TF)) {
                // candidate values from HBM, every pool entry at once (two
                // per lane, loads in flight together) instead of one memory
                // round trip per emit round
                static_assert(kPool <= 2 * kWave, "two pool entries per lane");
                wave_sync();
                const int used = uniform(*pool_n) < kPool ? uniform(*pool_n) : kPool;
                int e[2];
                bool has[2];
                double v[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    e[h] = lane + h * kWave;
                    const int ql = e[h] < used ? (int)pq[e[h]] : 0xff;
                    has[h] = ql != 0xff;
                    const uint32_t ij = has[h] ? pij[e[h]] : 0u;
                    const int i = (int)(ij >> 16), j = (int)(ij & 0xffffu);
                    if constexpr (F32)
                        v[h] = has[h] ? R.inp[roff + i] * Q.inp[qoff_l[ql] + j] : 0.0;
                    else
                        v[h] = has[h] ? pair_value(R, roff + i, Q, qoff_l[ql] + j) : 0.0;
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (has[h]) pv[e[h]] = v[h];
            }
        };
        auto queue_heavy = [&](int n) {
            const unsigned long long e = pair_entry(a.batched, a.job, r, q);
            const int64_t half = a.heavy_cap >> 1;
            if (n <= kWaveKeys && nr <= kWavePeaks &&
                Q.off[q + 1] - Q.off[q] <= kWavePeaks) {  // one wave per pair
                const unsigned long long slot = atomicAdd(&a.counters[kSmallCount], 1ull);
                if ((int64_t)slot < a.heavy_cap - half) a.heavy_queue[half + slot] = e;
            } else {  // a workgroup per pair: the region of its size class
                const int k = heavy_class(n);
                const int64_t region = half / kHeavyClasses;
                const unsigned long long slot = atomicAdd(&a.counters[kClassCount + k], 1ull);
                if ((int64_t)slot < region) a.heavy_queue[k * region + slot] = e;
            }
            out_score = 0.0f;
            out_matches = -1;
            out_overflow = 0;
        };

        // Pass 0 probes for every query.  A query whose candidates did not all
        // fit the pool (pdrop) but are at most kMidCap is probed again in a
        // later pass with a group of such queries: groups are cut from the
        // exclusive prefix sum of their counts in steps of kMidCap, so one
        // group's candidates fill at most 2 * kMidCap = kPool entries.
        static_assert(2 * kMidCap <= kPool, "a group of pending queries fits the pool");
        const double rnorm = c.norm;
        int grp = -1, ngroups = 0;
        // settles the pooled pairs of the queries in `filter` (pass 0: all)
        auto settle_pass = [&](unsigned long long filter, bool first) {
            // 2. settle: lane l = pair (r, query l); every valid lane ends with
            // its pair's output in out_*
            const int used = uniform(*pool_n);
            const int n = cnt[lane];
            const bool mine = valid && ((filter >> lane) & 1ull);
            const bool dropped = pdrop[lane] != 0;
            if (first) {
                if (mine && n > kMidCap) queue_heavy(n);
                const bool pending = mine && dropped && n <= kMidCap;
                if (__ballot(pending)) {  // rare: only after the pool overflowed
                    int x = pending ? n : 0, incl = x;
                    for (int o = 1; o < kWave; o <<= 1) {
                        const int y = __shfl_up(incl, o);
                        if (lane >= o) incl += y;
                    }
                    grp = pending ? (incl - x) / kMidCap : -1;
                    int top = grp + 1;
                    for (int o = kWave / 2; o > 0; o >>= 1) {
                        const int y = __shfl_xor(top, o);
                        top = y > top ? y : top;
                    }
                    ngroups = uniform(top);
                }
            } else if (mine && dropped) {  // cannot happen (a group fits); stay exact
                queue_heavy(n);
            }
            const bool settle = mine && !dropped && n <= kMidCap;
            // 2a. up to K candidates: walk the pair's list into registers
            const bool small = settle && n <= K;
            // wave-wide max of the small counts, by a binary search of ballots
            // (no LDS round trip); a scalar, so the unrolled steps past it are
            // branched over
            int bound = 0, top = K;
            while (bound < top) {
                const int mid = (bound + top + 1) >> 1;
                if (__ballot(small && n >= mid))
                    bound = mid;
                else
                    top = mid - 1;
            }
            static_assert(K <= LL::ranks, "every small pair's candidates are indexed by rank");
            LaneCandidates<K> cand;
            cand.reset();
#pragma unroll
            for (int s = 0; s < K; ++s) {  // gathers through the rank index
                if (s >= bound) {  // uniform
                    if constexpr (kLaneEarlyExit<K>)
                        break;
                    else
                        continue;
                }
                if (small && s < n) {  // rank s arrives after ranks 0 .. s - 1
                    const int e = pidx[s * kTileQueries + lane];
                    cand.insert_rank(s, pv[e], pij[e]);
                }
            }
            if (small) {
                double sc = 0.0;
                int m = 0;
                cand.greedy(sc, m, bound);
                out_score = n ? (float)(sc / (rnorm * qnorm_l[lane])) : 0.0f;
                out_matches = m;
                out_overflow = n >= a.match_limit;
            }
            // 2b. K < n <= 64: the whole wave settles the pair from registers
            // (out of line: its registers would otherwise be held, and
            // spilled, across every step)
            unsigned long long mid = __ballot(settle && n > K);
            while (mid) {
                const int sl = __builtin_ctzll(mid);
                mid &= mid - 1ull;
                const MidResult res = mid_settle(pool_off, used < kPool ? used : kPool, sl);
                if (lane == sl) {
                    out_score = (float)(res.sc / (rnorm * qnorm_l[sl]));
                    out_matches = res.m;
                    out_overflow = n >= a.match_limit;
                }
            }
            wave_sync();
            cnt[lane] = 0;
            pdrop[lane] = 0;
            if (lane == 0) *pool_n = 0;
            wave_sync();
        };
        if (tile_f32)
            probe(std::true_type{}, std::true_type{}, ~0ull);
        else
            probe(std::false_type{}, std::true_type{}, ~0ull);
        wave_sync();
        flush();  // the previous reference's outputs
        out_row = r;
        settle_pass(~0ull, true);
        // later passes: one group of pending queries each, reference peaks
        // read again from memory
        for (int g = 0; g < ngroups; ++g) {
            const unsigned long long filter = __ballot(grp == g);
            if (!filter) continue;
            if (tile_f32)
                probe(std::true_type{}, std::false_type{}, filter);
            else
                probe(std::false_type{}, std::false_type{}, filter);
            wave_sync();
            settle_pass(filter, false);
        }
    };

    // One wait per step, before the next prefetch is issued: what is then in
    // flight is this reference's peaks (issued a whole step ago) and the
    // stores issued after the previous probe.
    RefChunk nxt;
    fetch(nxt, grab());
    for (;;) {
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        const RefChunk cur = nxt;
        if (cur.rr >= nrefs) break;
        fetch(nxt, grab());
        step(cur);
    }
    flush();
}

// One call: workgroup b is (reference block b % n_ref_blocks, tile b / n_ref_blocks).
template <int K, bool F32>
__global__ void __launch_bounds__(kLightThreads) k_light(LightArgs a) {
    light_body<K, F32>(a, (int64_t)blockIdx.x % a.n_ref_blocks,
                       (int64_t)blockIdx.x / a.n_ref_blocks);
}

// One call, scheduled: a persistent grid takes (tile, reference block) items
// from a counter, tiles in `order` (most peaks first, order_tiles), so the
// longest items start first and the last round is made of the shortest.
template <int K, bool F32>
__global__ void __launch_bounds__(kLightThreads)
k_light_sched(LightArgs a, const uint32_t *__restrict__ order, int64_t n_items,
              unsigned long long *next) {
    __shared__ unsigned long long item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(next, 1ull);
        __syncthreads();
        const int64_t it = uniform64((int64_t)item_s);
        __syncthreads();
        if (it >= n_items) break;
        light_body<K, F32>(a, it % a.n_ref_blocks, (int64_t)order[it / a.n_ref_blocks]);
        __syncthreads();  // the next item restages the LDS
    }
}

// A batch: a persistent grid takes work items (job << 48 | reference block
// << 24 | tile), costliest first, from a shared counter until none is left.
// The job table and item list are read-only and alias nothing the kernel
// writes (__restrict__): without that the compiler reloads job fields after
// every store, about 20% of the kernel's time.
template <int K, bool F32>
__global__ void __launch_bounds__(kLightThreads)
k_light_batch(const LightArgs *__restrict__ jobs, const unsigned long long *__restrict__ items,
              int64_t n_items, unsigned long long *next) {
    __shared__ unsigned long long item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(next, 1ull);
        __syncthreads();
        const unsigned long long it = (unsigned long long)uniform64((int64_t)item_s);
        __syncthreads();
        if ((int64_t)it >= n_items) break;
        const unsigned long long w = items[it];  // scalar load: `it` is uniform
        light_body<K, F32>(jobs[w >> 48], (int64_t)((w >> 24) & 0xffffffull),
                           (int64_t)(w & 0xffffffull));
        __syncthreads();  // the next item restages the LDS
    }
}

// ---------------------------------------------------------------------------
// Heavy kernel: one workgroup per pair with more than 64 candidates
// ---------------------------------------------------------------------------

constexpr int kHeavyThreads = 512;
constexpr int kScanChunk = 256;    // pairs one workgroup checks per round of the overflow scan
constexpr int kHeavyBuf = 4096;     // candidates one pair settles in a single pass
// The heavy buffer keeps entry x at hphys(x) = x + x / 16: the sort's register
// passes read 16 consecutive entries per thread, and the pad slot spreads a
// wave's accesses over the LDS banks.
constexpr int kHeavyBufPad = kHeavyBuf + kHeavyBuf / 16;
constexpr int kHeavyStage = 1024;   // peaks per spectrum staged in LDS
constexpr int kHeavyBlocks = 1024;
constexpr int kGiantGrid = 512;

struct HeavyArgs {
    SpecView R, Q;
    int64_t nR, nQ;
    float tolc, shift32;
    double tol, shift;
    int64_t match_limit;
    int round_cap;         // giant path: candidates settled per round (T)
    int buf_cap;           // giant path: next_pow2(2T + 64) entries
    int r_words, q_words;  // bitset words for the largest spectra
    int stage;             // both spectra fit kHeavyStage: staged in LDS
    int exact_from_stage;  // staged query m/z are the exact values (no shift, no float64)
    float *score;
    int32_t *matches;
    uint8_t *overflow;
    int64_t ld;
    const unsigned long long *heavy_queue;
    int64_t heavy_cap;
    unsigned long long *counters;  // the call's counters (kCounters of them, listed above)
    unsigned long long *giant_queue;
    int job;                       // this job's index in a batch
    int batched;                   // queue entries carry the job index
};

__device__ __forceinline__ bool bit_set(const uint32_t *b, int x) {
    return (b[x >> 5] >> (x & 31)) & 1u;
}

// Queues pair (r, q) of `a` for k_giant (called by one thread).
__device__ __forceinline__ void queue_giant(const HeavyArgs &a, int64_t r, int64_t q) {
    const unsigned long long slot = atomicAdd(&a.counters[2], 1ull);
    if ((int64_t)slot < a.heavy_cap) a.giant_queue[slot] = pair_entry(a.batched, a.job, r, q);
}

// Pad [n, next_pow2(n)) with -inf keys and sort the first next_pow2(n)
// entries descending by key_greater, with every thread of the block.
// Thread t takes compare-exchange pairs t, t + nt, ...; a wave's 64
// consecutive pairs always span one aligned block of 128 keys, so stages of
// stride <= 64 stay inside the wave and need no workgroup barrier between them.
__device__ void sort_desc(double *bv, uint32_t *bij, int n) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int e = n + tid; e < n2; e += nt) {
        bv[e] = -__builtin_huge_val();
        bij[e] = 0u;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1) {
        // blocks of `size` entries that lie wholly in the -inf padding stay
        // padding through this size's stages: their compare-exchanges are
        // skipped (only blocks up to the one holding entry n - 1 are sorted)
        const int live = ((n + size - 1) / size) * size;
        const int pairs = (live < n2 ? live : n2) >> 1;
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < pairs; t += nt) {
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
            const int next = stride > 1 ? stride >> 1 : size;  // the next stage's stride
            if (stride > kWave || next > kWave)
                __syncthreads();
            else
                wave_sync();
        }
    }
    __syncthreads();
}

__device__ __forceinline__ int hphys(int x) { return x + (x >> 4); }

// Compare-exchange of two register entries (a at the lower index).
__device__ __forceinline__ void reg_ce(double &va, uint32_t &ia, double &vb, uint32_t &ib,
                                       bool desc) {
    const bool sw = desc ? key_greater(vb, ib, va, ia) : key_greater(va, ia, vb, ib);
    const double tv = va;
    const uint32_t ti = ia;
    va = sw ? vb : va;
    ia = sw ? ib : ia;
    vb = sw ? tv : vb;
    ib = sw ? ti : ib;
}

// The stages of strides E/2 .. 1 of one bitonic size on thread t's entries
// t*E .. t*E + E - 1, in registers (initial: every size 2 .. E).  Threads
// whose entries all lie at or past `live` (padding) do nothing.
template <int E>
__device__ void heavy_reg_pass(double *bv, uint32_t *bij, int size, bool initial, int live) {
    const int x0 = threadIdx.x * E;
    if (x0 >= live) return;
    double v[E];
    uint32_t ij[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        v[e] = bv[hphys(x0 + e)];
        ij[e] = bij[hphys(x0 + e)];
    }
    if (initial) {
#pragma unroll
        for (int sz = 2; sz <= E; sz <<= 1) {
#pragma unroll
            for (int st = sz >> 1; st > 0; st >>= 1) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & st) continue;
                    reg_ce(v[e], ij[e], v[e + st], ij[e + st], ((x0 + e) & sz) == 0);
                }
            }
        }
    } else {
        const bool desc = (x0 & size) == 0;  // size >= 2E: the same for all E entries
#pragma unroll
        for (int st = E >> 1; st > 0; st >>= 1) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (e & st) continue;
                reg_ce(v[e], ij[e], v[e + st], ij[e + st], desc);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        bv[hphys(x0 + e)] = v[e];
        bij[hphys(x0 + e)] = ij[e];
    }
}

// One LDS stage (partner distance `stride`) of bitonic size `size` over the
// first `pairs` compare-exchange pairs.
__device__ __forceinline__ void heavy_lds_stage(double *bv, uint32_t *bij, int size, int stride,
                                                int pairs) {
    for (int t = threadIdx.x; t < pairs; t += kHeavyThreads) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const int pl = hphys(lo), ph = hphys(hi);
        const bool desc = (lo & size) == 0;
        const double va = bv[pl], vb = bv[ph];
        const uint32_t ia = bij[pl], ib = bij[ph];
        const bool sw = desc ? key_greater(vb, ib, va, ia) : key_greater(va, ia, vb, ib);
        if (sw) {
            bv[pl] = vb;
            bv[ph] = va;
            bij[pl] = ib;
            bij[ph] = ia;
        }
    }
}

// Two LDS stages at once, partner distances 2h then h (h >= 1) of bitonic
// size `size` >= 4h, over the first `quads` groups of four entries x, x + h,
// x + 2h, x + 3h (x with both stride bits clear): one load and at most one
// store of each entry for the two stages.  An entry is stored only if it
// moved, i.e. its ij changed: real entries have distinct ij (i, j < 65535)
// and every -inf padding entry is (-inf, 0xffffffff).
__device__ __forceinline__ void heavy_lds_stage2(double *bv, uint32_t *bij, int size, int h,
                                                 int quads) {
    for (int t = threadIdx.x; t < quads; t += kHeavyThreads) {
        const int lowb = t & (h - 1);
        const int x = lowb + 4 * (t - lowb);
        const bool desc = (x & size) == 0;
        int p[4];
        double v[4];
        uint32_t ij[4], ij0[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            p[k] = hphys(x + k * h);
            v[k] = bv[p[k]];
            ij[k] = ij0[k] = bij[p[k]];
        }
        reg_ce(v[0], ij[0], v[2], ij[2], desc);
        reg_ce(v[1], ij[1], v[3], ij[3], desc);
        reg_ce(v[0], ij[0], v[1], ij[1], desc);
        reg_ce(v[2], ij[2], v[3], ij[3], desc);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (ij[k] != ij0[k]) {
                bv[p[k]] = v[k];
                bij[p[k]] = ij[k];
            }
        }
    }
}

// Descending bitonic sort of the heavy buffer (entries at hphys) by the
// kHeavyThreads threads.  With n2 = kHeavyThreads * E entries (E >= 2), the
// strides below E of every size run in registers (one load and one store of
// each entry per size) and only the larger strides go through LDS, two
// stages per pass.  Blocks wholly in the -inf padding are skipped.
template <int E>
__device__ void heavy_sort_e(double *bv, uint32_t *bij, int n, int n2) {
    heavy_reg_pass<E>(bv, bij, 0, true, ((n + E - 1) / E) * E);
    __syncthreads();
    for (int size = 2 * E; size <= n2; size <<= 1) {
        const int live = ((n + size - 1) / size) * size;
        const int pairs = (live < n2 ? live : n2) >> 1;
        int stride = size >> 1;
        for (; stride >= 2 * E; stride >>= 2) {  // two strides per LDS pass
            heavy_lds_stage2(bv, bij, size, stride >> 1, pairs >> 1);
            __syncthreads();
        }
        if (stride >= E) {
            heavy_lds_stage(bv, bij, size, stride, pairs);
            __syncthreads();
        }
        heavy_reg_pass<E>(bv, bij, size, false, live < n2 ? live : n2);
        __syncthreads();
    }
}

// One wave sorts n2 = 64 * EPL entries of the heavy buffer descending, in
// registers: lane l holds entries EPL * l .. EPL * l + EPL - 1.  Strides below
// EPL are exchanges inside a lane; a stride s >= EPL pairs lanes at distance
// s / EPL (DPP, swizzle or bpermute moves).  No workgroup barrier: a small
// pair's sort costs one wave a few thousand instructions instead of the whole
// workgroup a barrier per stage.
template <int EPL, int SIZE, int STRIDE>
__device__ __forceinline__ void wave_heavy_stage(double (&v)[EPL], uint32_t (&ij)[EPL], int lane) {
    if constexpr (STRIDE >= EPL) {
        constexpr int D = STRIDE / EPL;
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int x = EPL * lane + k;
            const bool desc = (x & SIZE) == 0;
            const bool lower = (x & STRIDE) == 0;
            const double ov = xor_lane_double<D>(v[k]);
            const uint32_t oij = (uint32_t)xor_lane<D>((int)ij[k]);
            // the lower entry of a descending block keeps the greater key
            const bool take = (lower == desc) ? key_greater(ov, oij, v[k], ij[k])
                                              : key_greater(v[k], ij[k], ov, oij);
            v[k] = take ? ov : v[k];
            ij[k] = take ? oij : ij[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            if (k & STRIDE) continue;
            reg_ce(v[k], ij[k], v[k + STRIDE], ij[k + STRIDE], ((EPL * lane + k) & SIZE) == 0);
        }
    }
    if constexpr (STRIDE > 1) wave_heavy_stage<EPL, SIZE, STRIDE / 2>(v, ij, lane);
}
template <int EPL, int SIZE>
__device__ __forceinline__ void wave_heavy_sizes(double (&v)[EPL], uint32_t (&ij)[EPL], int lane) {
    wave_heavy_stage<EPL, SIZE, SIZE / 2>(v, ij, lane);
    if constexpr (SIZE < kWave * EPL) wave_heavy_sizes<EPL, SIZE * 2>(v, ij, lane);
}
template <int EPL>
__device__ __attribute__((noinline)) void wave_sort_heavy(double *bv, uint32_t *bij, int n) {
    const int lane = threadIdx.x & (kWave - 1);
    double v[EPL];
    uint32_t ij[EPL];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
        const int x = EPL * lane + k;
        v[k] = x < n ? bv[hphys(x)] : -__builtin_huge_val();
        ij[k] = x < n ? bij[hphys(x)] : 0xffffffffu;
    }
    wave_heavy_sizes<EPL, 2>(v, ij, lane);
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
        bv[hphys(EPL * lane + k)] = v[k];
        bij[hphys(EPL * lane + k)] = ij[k];
    }
    wave_sync();
}

// Sorts the heavy buffer's first n entries descending.  Padded to 128 or 256
// entries, wave 0 alone sorts in registers (the other waves return at once);
// larger buffers are sorted by the whole workgroup (for 512 entries one wave
// was no faster).
__device__ void sort_desc_heavy(double *bv, uint32_t *bij, int n) {
    const int tid = threadIdx.x;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    if (n2 >= 2 * kWave && n2 <= 4 * kWave) {  // uniform
        if (tid < kWave) {
            if (n2 == 2 * kWave)
                wave_sort_heavy<2>(bv, bij, n);
            else
                wave_sort_heavy<4>(bv, bij, n);
        }
        return;
    }
    for (int e = n + tid; e < n2; e += kHeavyThreads) {  // no real entry has this ij
        bv[hphys(e)] = -__builtin_huge_val();
        bij[hphys(e)] = 0xffffffffu;
    }
    __syncthreads();
    switch (n2 / kHeavyThreads) {
    case 16: heavy_sort_e<16>(bv, bij, n, n2); return;
    case 8: heavy_sort_e<8>(bv, bij, n, n2); return;
    case 4: heavy_sort_e<4>(bv, bij, n, n2); return;
    case 2: heavy_sort_e<2>(bv, bij, n, n2); return;
    default: break;
    }
    // n2 <= kHeavyThreads: every stage through LDS
    for (int size = 2; size <= n2; size <<= 1) {
        const int live = ((n + size - 1) / size) * size;
        const int pairs = (live < n2 ? live : n2) >> 1;
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            heavy_lds_stage(bv, bij, size, stride, pairs);
            __syncthreads();
        }
    }
}

// ---- 64-bit keys of the staged heavy path ----
// A staged pair has at most kHeavyStage = 1024 peaks per side, so (i, j) fits
// in 20 bits.  Its key is the top 44 bits of an order-preserving map of v
// (the unsigned order of the map is the float64 order of v; -0 counts as +0)
// followed by i and j.  Keys order as (v desc, i desc, j desc) does, except
// between distinct values that share the 44-bit prefix: after the sort, the
// caller compares every adjacent pair with equal prefixes exactly and falls
// back to the (v, ij) sort if any is out of order.  One 64-bit compare and
// two selects replace three compares and six selects per compare-exchange,
// and an entry takes 8 bytes instead of 12.  The padding key 0 lies below
// every real key.
constexpr int kKeyIjBits = 20;
constexpr unsigned long long kKeyPrefix = ~((1ull << kKeyIjBits) - 1ull);
static_assert(kHeavyStage <= 1024, "i and j of a staged pair fit 10 bits each");

__device__ __forceinline__ unsigned long long heavy_key(double v, int i, int j) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v + 0.0);
    const unsigned long long ord = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    return (ord & kKeyPrefix) | ((unsigned long long)i << 10) | (unsigned long long)j;
}
__device__ __forceinline__ int key_i(unsigned long long k) { return (int)((k >> 10) & 1023u); }
__device__ __forceinline__ int key_j(unsigned long long k) { return (int)(k & 1023u); }

// Compare-exchange of two keys (a at the lower index).
__device__ __forceinline__ void key_ce(unsigned long long &a, unsigned long long &b, bool desc) {
    const bool sw = desc ? (b > a) : (a > b);
    const unsigned long long t = a;
    a = sw ? b : a;
    b = sw ? t : b;
}

// heavy_reg_pass on keys.
template <int E>
__device__ void key_reg_pass(unsigned long long *kb, int size, bool initial, int live) {
    const int x0 = threadIdx.x * E;
    if (x0 >= live) return;
    unsigned long long k[E];
#pragma unroll
    for (int e = 0; e < E; ++e) k[e] = kb[hphys(x0 + e)];
    if (initial) {
#pragma unroll
        for (int sz = 2; sz <= E; sz <<= 1) {
#pragma unroll
            for (int st = sz >> 1; st > 0; st >>= 1) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & st) continue;
                    key_ce(k[e], k[e + st], ((x0 + e) & sz) == 0);
                }
            }
        }
    } else {
        const bool desc = (x0 & size) == 0;
#pragma unroll
        for (int st = E >> 1; st > 0; st >>= 1) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (e & st) continue;
                key_ce(k[e], k[e + st], desc);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) kb[hphys(x0 + e)] = k[e];
}

// heavy_lds_stage on keys.
__device__ __forceinline__ void key_lds_stage(unsigned long long *kb, int size, int stride,
                                              int pairs) {
    for (int t = threadIdx.x; t < pairs; t += kHeavyThreads) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const int pl = hphys(lo), ph = hphys(hi);
        const bool desc = (lo & size) == 0;
        const unsigned long long a = kb[pl], b = kb[ph];
        if (desc ? (b > a) : (a > b)) {
            kb[pl] = b;
            kb[ph] = a;
        }
    }
}

// heavy_lds_stage2 on keys: an entry is stored only if it moved (real keys
// are distinct, every padding key is 0).
__device__ __forceinline__ void key_lds_stage2(unsigned long long *kb, int size, int h,
                                               int quads) {
    for (int t = threadIdx.x; t < quads; t += kHeavyThreads) {
        const int lowb = t & (h - 1);
        const int x = lowb + 4 * (t - lowb);
        const bool desc = (x & size) == 0;
        int p[4];
        unsigned long long k[4], k0[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            p[u] = hphys(x + u * h);
            k[u] = k0[u] = kb[p[u]];
        }
        key_ce(k[0], k[2], desc);
        key_ce(k[1], k[3], desc);
        key_ce(k[0], k[1], desc);
        key_ce(k[2], k[3], desc);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k[u] != k0[u]) kb[p[u]] = k[u];
    }
}

template <int E>
__device__ void key_sort_e(unsigned long long *kb, int n, int n2) {
    key_reg_pass<E>(kb, 0, true, ((n + E - 1) / E) * E);
    __syncthreads();
    for (int size = 2 * E; size <= n2; size <<= 1) {
        const int live = ((n + size - 1) / size) * size;
        const int pairs = (live < n2 ? live : n2) >> 1;
        int stride = size >> 1;
        for (; stride >= 2 * E; stride >>= 2) {
            key_lds_stage2(kb, size, stride >> 1, pairs >> 1);
            __syncthreads();
        }
        if (stride >= E) {
            key_lds_stage(kb, size, stride, pairs);
            __syncthreads();
        }
        key_reg_pass<E>(kb, size, false, live < n2 ? live : n2);
        __syncthreads();
    }
}

// x from lane (lane ^ D) of the wave, for a 64-bit key.
template <int D>
__device__ __forceinline__ unsigned long long xor_lane_key(unsigned long long x) {
    const unsigned lo = (unsigned)xor_lane<D>((int)(unsigned)x);
    const unsigned hi = (unsigned)xor_lane<D>((int)(unsigned)(x >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// wave_heavy_stage on keys.
template <int EPL, int SIZE, int STRIDE>
__device__ __forceinline__ void wave_key_stage(unsigned long long (&k)[EPL], int lane) {
    if constexpr (STRIDE >= EPL) {
        constexpr int D = STRIDE / EPL;
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const int x = EPL * lane + e;
            const bool desc = (x & SIZE) == 0;
            const bool lower = (x & STRIDE) == 0;
            const unsigned long long o = xor_lane_key<D>(k[e]);
            // the lower entry of a descending block keeps the greater key
            const bool take = (lower == desc) ? (o > k[e]) : (k[e] > o);
            k[e] = take ? o : k[e];
        }
    } else {
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            if (e & STRIDE) continue;
            key_ce(k[e], k[e + STRIDE], ((EPL * lane + e) & SIZE) == 0);
        }
    }
    if constexpr (STRIDE > 1) wave_key_stage<EPL, SIZE, STRIDE / 2>(k, lane);
}
template <int EPL, int SIZE>
__device__ __forceinline__ void wave_key_sizes(unsigned long long (&k)[EPL], int lane) {
    wave_key_stage<EPL, SIZE, SIZE / 2>(k, lane);
    if constexpr (SIZE < kWave * EPL) wave_key_sizes<EPL, SIZE * 2>(k, lane);
}
// One wave sorts 64 * EPL keys descending in registers (wave_sort_heavy).
template <int EPL>
__device__ __attribute__((noinline)) void wave_sort_keys(unsigned long long *kb, int n) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long k[EPL];
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
        const int x = EPL * lane + e;
        k[e] = x < n ? kb[hphys(x)] : 0ull;
    }
    wave_key_sizes<EPL, 2>(k, lane);
#pragma unroll
    for (int e = 0; e < EPL; ++e) kb[hphys(EPL * lane + e)] = k[e];
    wave_sync();
}

// Sorts the first n keys descending (sort_desc_heavy on keys).  Padded to
// 128 or 256 keys, wave 0 alone sorts and the other waves return at once;
// otherwise the whole workgroup sorts and every thread returns after a
// barrier.
__device__ void sort_keys(unsigned long long *kb, int n) {
    const int tid = threadIdx.x;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    if (n2 >= 2 * kWave && n2 <= 4 * kWave) {  // uniform
        if (tid < kWave) {
            if (n2 == 2 * kWave)
                wave_sort_keys<2>(kb, n);
            else
                wave_sort_keys<4>(kb, n);
        }
        return;
    }
    for (int e = n + tid; e < n2; e += kHeavyThreads) kb[hphys(e)] = 0ull;
    __syncthreads();
    switch (n2 / kHeavyThreads) {
    case 16: key_sort_e<16>(kb, n, n2); return;
    case 8: key_sort_e<8>(kb, n, n2); return;
    case 4: key_sort_e<4>(kb, n, n2); return;
    case 2: key_sort_e<2>(kb, n, n2); return;
    default: break;
    }
    for (int size = 2; size <= n2; size <<= 1) {
        const int live = ((n + size - 1) / size) * size;
        const int pairs = (live < n2 ? live : n2) >> 1;
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            key_lds_stage(kb, size, stride, pairs);
            __syncthreads();
        }
    }
}

// Greedy over sorted entries [0, n), run by one wave: chunks of 64 in order;
// inside a chunk the first live candidate is accepted and every later one
// sharing its i or j is dropped, repeatedly, which is exactly the sequential
// greedy.  A chunk's accepted peaks are marked used after the chunk.  The
// accepted sum runs in acceptance order.
template <bool PAD>
__device__ void greedy_sorted(const double *bv, const uint32_t *bij, int n, uint32_t *used_r,
                              uint32_t *used_q, double &s, int &m) {
    const int lane = threadIdx.x & (kWave - 1);
    for (int base = 0; base < n; base += kWave) {
        const int idx = base + lane;
        const bool valid = idx < n;
        const int px = PAD ? hphys(idx) : idx;
        const uint32_t ij = valid ? bij[px] : 0u;
        const double v = valid ? bv[px] : 0.0;
        const int i = (int)(ij >> 16), j = (int)(ij & 0xffffu);
        bool alive = (int)valid & (int)!bit_set(used_r, i) & (int)!bit_set(used_q, j);
        bool acc = false;
        unsigned long long mask = __ballot(alive);
        while (mask) {
            const int f = __builtin_ctzll(mask);
            const int fi = lane_int(i, f), fj = lane_int(j, f);
            s += lane_double(v, f);
            ++m;
            acc = acc | (lane == f);
            alive = alive & (lane != f) & (i != fi) & (j != fj);
            mask = __ballot(alive);
        }
        if (acc) {
            atomicOr(&used_r[i >> 5], 1u << (i & 31));
            atomicOr(&used_q[j >> 5], 1u << (j & 31));
        }
        wave_sync();
    }
}

// The same greedy over the heavy buffer, a chunk of 64 in rounds: a round
// accepts every live candidate that no earlier live candidate of the chunk
// shares a peak with (the sequential greedy accepts it: what it accepts
// before it is live, and none of that clashes), then drops every live
// candidate that shares a peak with an accepted one.  Heavy pairs accept
// about 14 candidates per chunk in under 2 rounds.  own_r / own_q (one word
// per peak, all kWave on entry and on return) hold the earliest live lane of
// each peak during a round.  The accepted values are summed in chunk order,
// as the sequential greedy sums them.
__device__ void greedy_rounds(const double *bv, const uint32_t *bij, int n, uint32_t *used_r,
                              uint32_t *used_q, uint32_t *own_r, uint32_t *own_q, double &s,
                              int &m) {
    const int lane = threadIdx.x & (kWave - 1);
    for (int base = 0; base < n; base += kWave) {
        const int idx = base + lane;
        const bool valid = idx < n;
        const uint32_t ij = valid ? bij[hphys(idx)] : 0u;
        const double v = valid ? bv[hphys(idx)] : 0.0;
        const int i = (int)(ij >> 16), j = (int)(ij & 0xffffu);
        bool alive = (int)valid & (int)!bit_set(used_r, i) & (int)!bit_set(used_q, j);
        unsigned long long acc = 0ull;
        while (__ballot(alive)) {
            if (alive) {
                atomicMin(&own_r[i], (uint32_t)lane);
                atomicMin(&own_q[j], (uint32_t)lane);
            }
            wave_sync();
            const uint32_t er = alive ? own_r[i] : (uint32_t)kWave;
            const uint32_t eq = alive ? own_q[j] : (uint32_t)kWave;
            const bool take = alive & (er == (uint32_t)lane) & (eq == (uint32_t)lane);
            const unsigned long long f = __ballot(take);
            acc |= f;
            // a live lane whose earliest same-peak lane was taken loses that peak
            const bool lost = alive & ((((f >> (er & 63u)) | (f >> (eq & 63u))) & 1ull) != 0ull);
            wave_sync();
            if (alive) {
                own_r[i] = (uint32_t)kWave;
                own_q[j] = (uint32_t)kWave;
            }
            wave_sync();
            alive = alive & !take & !lost;
        }
        for (unsigned long long a = acc; a; a &= a - 1ull) s += lane_double(v, __builtin_ctzll(a));
        m += __popcll(acc);
        if ((acc >> lane) & 1ull) {
            atomicOr(&used_r[i >> 5], 1u << (i & 31));
            atomicOr(&used_q[j >> 5], 1u << (j & 31));
        }
        wave_sync();
    }
}

// greedy_rounds over sorted keys: the same rounds on (i, j) decoded from the
// keys; only the accepted candidates' values are computed (value(i, j), the
// enumeration's own product), and they are summed in chunk order as there.
template <class Value>
__device__ void greedy_keys(const unsigned long long *kb, int n, uint32_t *used_r,
                            uint32_t *used_q, uint32_t *own_r, uint32_t *own_q, Value value,
                            double &s, int &m) {
    const int lane = threadIdx.x & (kWave - 1);
    for (int base = 0; base < n; base += kWave) {
        const int idx = base + lane;
        const bool valid = idx < n;
        const unsigned long long k = valid ? kb[hphys(idx)] : 0ull;
        const int i = key_i(k), j = key_j(k);
        bool alive = (int)valid & (int)!bit_set(used_r, i) & (int)!bit_set(used_q, j);
        unsigned long long acc = 0ull;
        while (__ballot(alive)) {
            if (alive) {
                atomicMin(&own_r[i], (uint32_t)lane);
                atomicMin(&own_q[j], (uint32_t)lane);
            }
            wave_sync();
            const uint32_t er = alive ? own_r[i] : (uint32_t)kWave;
            const uint32_t eq = alive ? own_q[j] : (uint32_t)kWave;
            const bool take = alive & (er == (uint32_t)lane) & (eq == (uint32_t)lane);
            const unsigned long long f = __ballot(take);
            acc |= f;
            const bool lost = alive & ((((f >> (er & 63u)) | (f >> (eq & 63u))) & 1ull) != 0ull);
            wave_sync();
            if (alive) {
                own_r[i] = (uint32_t)kWave;
                own_q[j] = (uint32_t)kWave;
            }
            wave_sync();
            alive = alive & !take & !lost;
        }
        const bool mine = (acc >> lane) & 1ull;
        const double v = mine ? value(i, j) : 0.0;
        for (unsigned long long t = acc; t; t &= t - 1ull) s += lane_double(v, __builtin_ctzll(t));
        m += __popcll(acc);
        if (mine) {
            atomicOr(&used_r[i >> 5], 1u << (i & 31));
            atomicOr(&used_q[j >> 5], 1u << (j & 31));
        }
        wave_sync();
    }
}

__device__ void write_error(const HeavyArgs &a, int64_t o) {
    a.score[o] = __builtin_nanf("");
    a.matches[o] = -2;
    a.overflow[o] = 1;
}

// One pair, whole block, by (v, ij) entries: stage both spectra, enumerate
// every candidate once into LDS, sort, greedy.  More than kHeavyBuf
// candidates: queued as giant.  The path of launches whose spectra are not
// staged (heavy_pair_keys is the staged one).
__device__ __forceinline__ void heavy_pair(const HeavyArgs &a, int64_t r, int64_t q,
                                           unsigned char *lds, int *bn_s) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const SpecView &R = a.R, &Q = a.Q;
    const int64_t roff = R.off[r], qoff = Q.off[q];
    const int nr = (int)(R.off[r + 1] - roff), nq = (int)(Q.off[q + 1] - qoff);
    const int64_t o = r * a.ld + q;
    const int S = a.stage ? kHeavyStage : 0;
    double *bv = reinterpret_cast<double *>(lds);
    double *r_w = bv + kHeavyBufPad;
    double *q_d = r_w + S;
    uint32_t *bij = reinterpret_cast<uint32_t *>(q_d + S);
    float *r_mz = reinterpret_cast<float *>(bij + kHeavyBufPad);
    float *q_ms = r_mz + S;
    uint32_t *used_r = reinterpret_cast<uint32_t *>(q_ms + S);
    uint32_t *used_q = used_r + a.r_words;
    const int nrw = (nr + 31) >> 5, nqw = (nq + 31) >> 5;
    if (nrw > a.r_words || nqw > a.q_words || (a.stage && (nr > S || nq > S))) {
        if (tid == 0) write_error(a, o);  // caller's max_peaks was wrong
        return;
    }
    if (a.stage) {
        for (int k = tid; k < nr; k += nt) {
            r_mz[k] = R.mz[roff + k];
            r_w[k] = (R.mzp ? R.mzp[roff + k] : 1.0) * R.inp[roff + k];
        }
        for (int k = tid; k < nq; k += nt) {
            q_ms[k] = Q.mz[qoff + k] + a.shift32;
            q_d[k] = Q.inp[qoff + k];
        }
    }
    for (int w = tid; w < nrw; w += nt) used_r[w] = 0u;
    for (int w = tid; w < nqw; w += nt) used_q[w] = 0u;
    if (tid == 0) *bn_s = 0;
    __syncthreads();

    auto rmz = [&](int i) { return a.stage ? r_mz[i] : R.mz[roff + i]; };
    auto qms = [&](int j) { return a.stage ? q_ms[j] : Q.mz[qoff + j] + a.shift32; };
    const int lane = tid & (kWave - 1);
    for (int ib = tid - lane; ib < nr; ib += nt) {  // wave-uniform rounds
        const int i = ib + lane;
        bool act = i < nr;
        const float ri = act ? rmz(i) : 0.0f;
        int j = nq;
        if (act) {  // first query peak not above the window
            int lo = 0, hi = nq;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ri - qms(mid) > a.tolc) lo = mid + 1;
                else hi = mid;
            }
            j = lo;
        }
        act = act && j < nq;
        const double mr = act ? (R.mzx ? R.mzx[roff + i] : (double)ri) : 0.0;
        const double wr =
            act ? (a.stage ? r_w[i] : (R.mzp ? R.mzp[roff + i] : 1.0) * R.inp[roff + i]) : 0.0;
        // lockstep over the lanes' windows: one candidate per lane and step,
        // one LDS atomic per wave and step
        while (__ballot(act)) {
            bool push = false;
            double v = 0.0;
            uint32_t ij = 0u;
            if (act) {
                const float qj = qms(j);
                if (ri - qj < -a.tolc) {
                    act = false;
                } else {
                    const double mq = a.exact_from_stage ? (double)qj : exact_mz(Q, qoff + j);
                    if (exact_window(mr, mq, a.tol, a.shift)) {
                        v = wr * (Q.mzp ? Q.mzp[qoff + j] : 1.0);
                        v = v * (a.stage ? q_d[j] : Q.inp[qoff + j]);
                        ij = pack_ij(i, j);
                        push = true;
                    }
                    ++j;
                    act = j < nq;
                }
            }
            const unsigned long long mask = __ballot(push);
            int base = 0;
            if (lane == 0 && mask) base = atomicAdd(bn_s, __popcll(mask));
            base = __shfl(base, 0);
            if (push) {
                const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
                if (pos < kHeavyBuf) {
                    bv[hphys(pos)] = v;
                    bij[hphys(pos)] = ij;
                }
            }
        }
    }
    __syncthreads();
    const int C = *bn_s;
    if (C > kHeavyBuf) {
        if (tid == 0) queue_giant(a, r, q);
        __syncthreads();
        return;
    }
    sort_desc_heavy(bv, bij, C);
    if (tid < kWave) {
        double sc = 0.0;
        int m = 0;
        if (a.stage) {  // the staged spectra are no longer read: their space holds own_r / own_q
            uint32_t *own_r = reinterpret_cast<uint32_t *>(r_w);
            uint32_t *own_q = own_r + kHeavyStage;
            for (int k = tid; k < nr; k += kWave) own_r[k] = (uint32_t)kWave;
            for (int k = tid; k < nq; k += kWave) own_q[k] = (uint32_t)kWave;
            wave_sync();
            greedy_rounds(bv, bij, C, used_r, used_q, own_r, own_q, sc, m);
        } else {
            greedy_sorted<true>(bv, bij, C, used_r, used_q, sc, m);
        }
        if (tid == 0) {
            a.score[o] = (float)(sc / (R.norm[r] * Q.norm[q]));
            a.matches[o] = m;
            a.overflow[o] = (uint8_t)(C >= a.match_limit);
        }
    }
    __syncthreads();
}

// One pair of a staged launch, whole block, by 64-bit keys (heavy_key):
// stage both spectra, enumerate every candidate once into LDS, sort the keys,
// check the order of adjacent keys with equal prefixes exactly, greedy.  More
// than kHeavyBuf candidates, or keys that left two distinct values out of
// order: queued as giant (k_giant settles any pair exactly, by (v, ij)).
__device__ __forceinline__ void heavy_pair_keys(const HeavyArgs &a, int64_t r, int64_t q,
                                                unsigned char *lds, int *bn_s) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const SpecView &R = a.R, &Q = a.Q;
    const int64_t roff = R.off[r], qoff = Q.off[q];
    const int nr = (int)(R.off[r + 1] - roff), nq = (int)(Q.off[q + 1] - qoff);
    const int64_t o = r * a.ld + q;
    constexpr int S = kHeavyStage;
    // heavy_pair_exact's layout; the keys lie over its value array, and its
    // ij array (not used here) holds own_r / own_q for the greedy
    unsigned long long *kb = reinterpret_cast<unsigned long long *>(lds);
    double *r_w = reinterpret_cast<double *>(lds) + kHeavyBufPad;
    double *q_d = r_w + S;
    uint32_t *own_r = reinterpret_cast<uint32_t *>(q_d + S);
    uint32_t *own_q = own_r + S;
    float *r_mz = reinterpret_cast<float *>(own_r + kHeavyBufPad);
    float *q_ms = r_mz + S;
    uint32_t *used_r = reinterpret_cast<uint32_t *>(q_ms + S);
    uint32_t *used_q = used_r + a.r_words;
    const int nrw = (nr + 31) >> 5, nqw = (nq + 31) >> 5;
    if (nrw > a.r_words || nqw > a.q_words || nr > S || nq > S) {
        if (tid == 0) write_error(a, o);  // caller's max_peaks was wrong
        return;
    }
    for (int k = tid; k < nr; k += nt) {
        r_mz[k] = R.mz[roff + k];
        r_w[k] = (R.mzp ? R.mzp[roff + k] : 1.0) * R.inp[roff + k];
        own_r[k] = (uint32_t)kWave;
    }
    for (int k = tid; k < nq; k += nt) {
        q_ms[k] = Q.mz[qoff + k] + a.shift32;
        q_d[k] = Q.inp[qoff + k];
        own_q[k] = (uint32_t)kWave;
    }
    for (int w = tid; w < nrw; w += nt) used_r[w] = 0u;
    for (int w = tid; w < nqw; w += nt) used_q[w] = 0u;
    if (tid == 0) *bn_s = 0;
    __syncthreads();

    // a candidate's value, as the enumeration computes it
    auto value = [&](int i, int j) {
        const double v = r_w[i] * (Q.mzp ? Q.mzp[qoff + j] : 1.0);
        return v * q_d[j];
    };
    const int lane = tid & (kWave - 1);
    for (int ib = tid - lane; ib < nr; ib += nt) {  // wave-uniform rounds
        const int i = ib + lane;
        bool act = i < nr;
        const float ri = act ? r_mz[i] : 0.0f;
        int j = nq;
        if (act) {  // first query peak not above the window
            int lo = 0, hi = nq;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ri - q_ms[mid] > a.tolc) lo = mid + 1;
                else hi = mid;
            }
            j = lo;
        }
        act = act && j < nq;
        const double mr = act ? (R.mzx ? R.mzx[roff + i] : (double)ri) : 0.0;
        // lockstep over the lanes' windows: one candidate per lane and step,
        // one LDS atomic per wave and step
        while (__ballot(act)) {
            bool push = false;
            unsigned long long key = 0ull;
            if (act) {
                const float qj = q_ms[j];
                if (ri - qj < -a.tolc) {
                    act = false;
                } else {
                    const double mq = a.exact_from_stage ? (double)qj : exact_mz(Q, qoff + j);
                    if (exact_window(mr, mq, a.tol, a.shift)) {
                        key = heavy_key(value(i, j), i, j);
                        push = true;
                    }
                    ++j;
                    act = j < nq;
                }
            }
            const unsigned long long mask = __ballot(push);
            int base = 0;
            if (lane == 0 && mask) base = atomicAdd(bn_s, __popcll(mask));
            base = __shfl(base, 0);
            if (push) {
                const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
                if (pos < kHeavyBuf) kb[hphys(pos)] = key;
            }
        }
    }
    __syncthreads();
    const int C = *bn_s;
    auto to_giant = [&]() {
        if (tid == 0) queue_giant(a, r, q);
        __syncthreads();
    };
    if (C > kHeavyBuf) {
        to_giant();
        return;
    }
    sort_keys(kb, C);
    // Adjacent keys with equal prefixes, in exact order?  A buffer sorted by
    // wave 0 alone is checked by wave 0 (the other waves returned at once); a
    // workgroup sort ended with a barrier.
    const bool by_wave = C > kWave && C <= 4 * kWave;
    int bad = 0;
    if (!by_wave || tid < kWave) {
        for (int x = tid; x + 1 < C; x += by_wave ? kWave : nt) {
            const unsigned long long k0 = kb[hphys(x)], k1 = kb[hphys(x + 1)];
            if (((k0 ^ k1) & kKeyPrefix) == 0ull) {
                const int i0 = key_i(k0), j0 = key_j(k0), i1 = key_i(k1), j1 = key_j(k1);
                bad |= (int)key_greater(value(i1, j1), pack_ij(i1, j1), value(i0, j0),
                                        pack_ij(i0, j0));
            }
        }
    }
    if (__syncthreads_or(bad)) {  // the sentinel -1 is still in the output
        to_giant();
        return;
    }
    if (tid < kWave) {
        double sc = 0.0;
        int m = 0;
        greedy_keys(kb, C, used_r, used_q, own_r, own_q, value, sc, m);
        if (tid == 0) {
            a.score[o] = (float)(sc / (R.norm[r] * Q.norm[q]));
            a.matches[o] = m;
            a.overflow[o] = (uint8_t)(C >= a.match_limit);
        }
    }
    __syncthreads();
}

// Bytes of LDS one wave uses for a small pair (heavy_pair_wave): keys at
// hphys positions, weighted intensities of both spectra, their m/z (later
// the greedy's owner slots), and the used bitsets.
constexpr int kWaveKeySlots = (kWaveKeys + kWaveKeys / 16 + 15) / 16 * 16;
constexpr size_t kWavePairBytes = (size_t)kWaveKeySlots * 8 + (size_t)kWavePeaks * (8 + 8 + 4 + 4) +
                                  2 * (size_t)(kWavePeaks / 32) * 4;
static_assert(kHeavyThreads / kWave * kWavePairBytes <= (size_t)(kHeavyBufPad + 2 * kHeavyStage) * 12,
              "every wave of a staged heavy workgroup has room for a small pair");
static_assert(kWaveKeys <= 4 * kWave, "a small pair's keys are sorted by one wave");

// One small pair (at most kWaveKeys candidates, spectra of at most kWavePeaks
// peaks) settled by one wave in its own LDS slice: the steps of
// heavy_pair_keys (staging, enumeration, key sort, exact check of equal
// prefixes, greedy) with wave-local synchronisation only, so the result is
// the same bit for bit.  Returns false, with the output untouched, for a pair
// that does not fit (the caller queues it as giant).  Every early return
// is on a value moved to a scalar register: the wave leaves together, and
// the caller's loop stays uniform.
__device__ bool heavy_pair_wave(const HeavyArgs &a, int64_t r, int64_t q, unsigned char *slice) {
    const int lane = threadIdx.x & (kWave - 1);
    const SpecView &R = a.R, &Q = a.Q;
    const int64_t roff = uniform64(R.off[r]), qoff = uniform64(Q.off[q]);
    const int nr = uniform((int)(R.off[r + 1] - roff)), nq = uniform((int)(Q.off[q + 1] - qoff));
    const int64_t o = r * a.ld + q;
    unsigned long long *kb = reinterpret_cast<unsigned long long *>(slice);
    double *r_w = reinterpret_cast<double *>(kb + kWaveKeySlots);
    double *q_d = r_w + kWavePeaks;
    float *r_mz = reinterpret_cast<float *>(q_d + kWavePeaks);
    float *q_ms = r_mz + kWavePeaks;
    uint32_t *own_r = reinterpret_cast<uint32_t *>(r_mz);  // after the enumeration
    uint32_t *own_q = reinterpret_cast<uint32_t *>(q_ms);
    uint32_t *used_r = reinterpret_cast<uint32_t *>(q_ms + kWavePeaks);
    uint32_t *used_q = used_r + kWavePeaks / 32;
    if (nr > kWavePeaks || nq > kWavePeaks) return false;
    for (int k = lane; k < nr; k += kWave) {
        r_mz[k] = R.mz[roff + k];
        r_w[k] = (R.mzp ? R.mzp[roff + k] : 1.0) * R.inp[roff + k];
    }
    for (int k = lane; k < nq; k += kWave) {
        q_ms[k] = Q.mz[qoff + k] + a.shift32;
        q_d[k] = Q.inp[qoff + k];
    }
    if (lane < kWavePeaks / 32) {
        used_r[lane] = 0u;
        used_q[lane] = 0u;
    }
    wave_sync();

    auto value = [&](int i, int j) {  // as heavy_pair_keys computes it
        const double v = r_w[i] * (Q.mzp ? Q.mzp[qoff + j] : 1.0);
        return v * q_d[j];
    };
    int C = 0;  // uniform
    for (int ib = 0; ib < nr; ib += kWave) {
        const int i = ib + lane;
        bool act = i < nr;
        const float ri = act ? r_mz[i] : 0.0f;
        int j = nq;
        if (act) {  // first query peak not above the window
            int lo = 0, hi = nq;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ri - q_ms[mid] > a.tolc) lo = mid + 1;
                else hi = mid;
            }
            j = lo;
        }
        act = act && j < nq;
        const double mr = act ? (R.mzx ? R.mzx[roff + i] : (double)ri) : 0.0;
        while (__ballot(act)) {
            bool push = false;
            unsigned long long key = 0ull;
            if (act) {
                const float qj = q_ms[j];
                if (ri - qj < -a.tolc) {
                    act = false;
                } else {
                    const double mq = a.exact_from_stage ? (double)qj : exact_mz(Q, qoff + j);
                    if (exact_window(mr, mq, a.tol, a.shift)) {
                        key = heavy_key(value(i, j), i, j);
                        push = true;
                    }
                    ++j;
                    act = j < nq;
                }
            }
            const unsigned long long mask = __ballot(push);
            if (push) {
                const int pos = C + __popcll(mask & ((1ull << lane) - 1ull));
                if (pos < kWaveKeys) kb[hphys(pos)] = key;
            }
            C += __popcll(mask);
        }
    }
    wave_sync();
    C = uniform(C);
    if (C > kWaveKeys) return false;
    if (C <= 2 * kWave)
        wave_sort_keys<2>(kb, C);
    else
        wave_sort_keys<4>(kb, C);
    int bad = 0;
    for (int x = lane; x + 1 < C; x += kWave) {
        const unsigned long long k0 = kb[hphys(x)], k1 = kb[hphys(x + 1)];
        if (((k0 ^ k1) & kKeyPrefix) == 0ull) {
            const int i0 = key_i(k0), j0 = key_j(k0), i1 = key_i(k1), j1 = key_j(k1);
            bad |= (int)key_greater(value(i1, j1), pack_ij(i1, j1), value(i0, j0), pack_ij(i0, j0));
        }
    }
    if (__ballot(bad != 0)) return false;  // the sentinel -1 is still in the output
    for (int k = lane; k < nr; k += kWave) own_r[k] = (uint32_t)kWave;
    for (int k = lane; k < nq; k += kWave) own_q[k] = (uint32_t)kWave;
    wave_sync();
    double sc = 0.0;
    int m = 0;
    greedy_keys(kb, C, used_r, used_q, own_r, own_q, value, sc, m);
    if (lane == 0) {
        a.score[o] = (float)(sc / (R.norm[r] * Q.norm[q]));
        a.matches[o] = m;
        a.overflow[o] = (uint8_t)(C >= a.match_limit);
    }
    wave_sync();
    return true;
}

// Settles the queued pairs: one workgroup per pair, the largest first.  A
// batch passes its job table (the queue and counters are shared).  When the
// queue overflowed, every job's matrix is scanned for the -1 sentinel.
// STAGED launches (both spectra of every pair staged in LDS) sort 64-bit keys
// (heavy_pair_keys), the others (v, ij) entries (heavy_pair); one path per
// instantiation keeps each within the register budget.
template <bool STAGED>
__device__ __forceinline__ void heavy_body(const HeavyArgs &a, const HeavyArgs *jobs, int n_jobs) {
    extern __shared__ __align__(16) unsigned char heavy_lds[];
    __shared__ int bn_s;
    __shared__ unsigned long long item_s;
    __shared__ int nlist;
    __shared__ long long list[kScanChunk];
    // The small pairs (the queue's upper half) are settled one per wave
    // first; then the workgroup pairs, largest size class first.
    const int64_t half = a.heavy_cap >> 1, region = half / kHeavyClasses;
    unsigned long long cls[kHeavyClasses], count = 0;
    bool fits = true;
#pragma unroll
    for (int k = 0; k < kHeavyClasses; ++k) {
        cls[k] = a.counters[kClassCount + k];
        count += cls[k];
        fits = fits && (int64_t)cls[k] <= region;
    }
    const unsigned long long small = a.counters[kSmallCount];
    if (count + small == 0) return;
    const int tid = threadIdx.x;
    if (fits && (int64_t)small <= a.heavy_cap - half) {
        if constexpr (STAGED) {
            // The small pairs first, one per wave: every resident wave takes
            // at most a few, so they form no tail.  One claim per workgroup
            // for all its waves (one atomic on the shared counter, not one
            // per wave).  The wave's item and pair are in scalar registers,
            // so its branches are uniform and it runs each pair whole.
            const int wave = uniform(tid >> 6);
            unsigned char *slice = heavy_lds + wave * kWavePairBytes;
            constexpr int kWaves = kHeavyThreads / kWave;
            for (;;) {
                if (tid == 0) item_s = atomicAdd(&a.counters[kSmallNext], (unsigned long long)kWaves);
                __syncthreads();
                const unsigned long long base = (unsigned long long)uniform64((int64_t)item_s);
                __syncthreads();
                if (base >= small) break;  // the slices become the workgroup's buffers
                const unsigned long long item = base + (unsigned long long)wave;
                if (item < small) {
                    const PairRef p =
                        pair_of(a.batched, uniform64((int64_t)a.heavy_queue[half + item]));
                    const HeavyArgs &J = jobs ? jobs[p.job] : a;
                    if (!heavy_pair_wave(J, p.r, p.q, slice) && (tid & (kWave - 1)) == 0)
                        queue_giant(J, p.r, p.q);
                }
            }
        }
        for (;;) {
            if (tid == 0) item_s = atomicAdd(&a.counters[1], 1ull);
            __syncthreads();
            const unsigned long long item = (unsigned long long)uniform64((int64_t)item_s);
            __syncthreads();
            if (item >= count) break;
            // the item's class and place in it, classes 3 .. 0 in turn
            static_assert(kHeavyClasses == 4, "four size classes");
            unsigned long long idx = item;
            int k;
            if (idx < cls[3]) {
                k = 3;
            } else if ((idx -= cls[3]) < cls[2]) {
                k = 2;
            } else if ((idx -= cls[2]) < cls[1]) {
                k = 1;
            } else {
                idx -= cls[1];
                k = 0;
            }
            const PairRef p = pair_of(a.batched, a.heavy_queue[k * region + (int64_t)idx]);
            if constexpr (STAGED)
                heavy_pair_keys(jobs ? jobs[p.job] : a, p.r, p.q, heavy_lds, &bn_s);
            else
                heavy_pair(jobs ? jobs[p.job] : a, p.r, p.q, heavy_lds, &bn_s);
        }
        if constexpr (!STAGED) {  // spectra too large to stage: a workgroup per small pair
            for (;;) {
                if (tid == 0) item_s = atomicAdd(&a.counters[kSmallNext], 1ull);
                __syncthreads();
                const unsigned long long item = (unsigned long long)uniform64((int64_t)item_s);
                __syncthreads();
                if (item >= small) break;
                const PairRef p = pair_of(a.batched, a.heavy_queue[half + item]);
                heavy_pair(jobs ? jobs[p.job] : a, p.r, p.q, heavy_lds, &bn_s);
            }
        }
    } else {  // queue overflowed: find the -1 sentinels
        for (int jb = 0; jb < n_jobs; ++jb) {
            const HeavyArgs &J = jobs ? jobs[jb] : a;
            const int64_t total = J.nR * J.nQ;
            for (int64_t c = blockIdx.x; c * kScanChunk < total; c += gridDim.x) {
                if (tid == 0) nlist = 0;
                __syncthreads();
                const int64_t p = c * kScanChunk + tid;
                if (tid < kScanChunk && p < total) {
                    const int64_t r = p / J.nQ, q = p - r * J.nQ;
                    if (J.matches[r * J.ld + q] == -1) list[atomicAdd(&nlist, 1)] = p;
                }
                __syncthreads();
                const int n = nlist;
                for (int k = 0; k < n; ++k) {
                    const int64_t pp = list[k];
                    if constexpr (STAGED)
                        heavy_pair_keys(J, pp / J.nQ, pp % J.nQ, heavy_lds, &bn_s);
                    else
                        heavy_pair(J, pp / J.nQ, pp % J.nQ, heavy_lds, &bn_s);
                }
                __syncthreads();
            }
        }
    }
}

// Two workgroups per compute unit (the LDS allows two): at most 128 VGPRs.
template <bool STAGED>
__global__ void __launch_bounds__(kHeavyThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_heavy(HeavyArgs a) {
    heavy_body<STAGED>(a, nullptr, 1);
}

template <bool STAGED>
__global__ void __launch_bounds__(kHeavyThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_heavy_batch(const HeavyArgs *__restrict__ jobs, int n_jobs) {
    heavy_body<STAGED>(jobs[0], jobs, n_jobs);
}

// Giant pairs (more than kHeavyBuf candidates): one wave per pair, in rounds.
// A round collects the best T alive candidates below the previous round's last
// key (a streaming top-T with a rising cut), settles them in order, and
// continues, so LDS stays bounded for any candidate count and the result is
// exactly the sequential greedy.
__device__ void giant_pair(const HeavyArgs &a, int64_t r, int64_t q, uint32_t *used_r,
                           uint32_t *used_q, double *bv, uint32_t *bij) {
    const int lane = threadIdx.x;
    const int64_t roff = a.R.off[r], qoff = a.Q.off[q];
    const int nr = (int)(a.R.off[r + 1] - roff), nq = (int)(a.Q.off[q + 1] - qoff);
    const int nrw = (nr + 31) >> 5, nqw = (nq + 31) >> 5;
    const int64_t o = r * a.ld + q;
    if (nrw > a.r_words || nqw > a.q_words) {
        if (lane == 0) write_error(a, o);
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
            if (act) {
                int lo = 0, hi = nq;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (ri - (qmz[mid] + a.shift32) > a.tolc) lo = mid + 1;
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
        greedy_sorted<false>(bv, bij, take, used_r, used_q, s, m);
        __syncthreads();
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

__device__ void giant_body(const HeavyArgs &a, const HeavyArgs *jobs, int n_jobs) {
    extern __shared__ __align__(16) unsigned char giant_lds[];
    double *bv = reinterpret_cast<double *>(giant_lds);
    uint32_t *bij = reinterpret_cast<uint32_t *>(bv + a.buf_cap);
    uint32_t *used_r = bij + a.buf_cap;
    uint32_t *used_q = used_r + a.r_words;

    const unsigned long long count = a.counters[2];
    if (count == 0) return;
    const int lane = threadIdx.x;
    if ((int64_t)count <= a.heavy_cap) {
        for (;;) {
            unsigned long long item = 0;
            if (lane == 0) item = atomicAdd(&a.counters[3], 1ull);
            item = (unsigned long long)uniform64((int64_t)__shfl(item, 0));
            if (item >= count) break;
            const PairRef p = pair_of(a.batched, a.giant_queue[item]);
            giant_pair(jobs ? jobs[p.job] : a, p.r, p.q, used_r, used_q, bv, bij);
        }
    } else {  // after k_heavy, the pairs still at -1 are exactly the giant ones
        for (int jb = 0; jb < n_jobs; ++jb) {
            const HeavyArgs &J = jobs ? jobs[jb] : a;
            const int64_t total = J.nR * J.nQ;
            for (int64_t c = blockIdx.x; c * kWave < total; c += gridDim.x) {
                const int64_t p = c * kWave + lane;
                int64_t r = 0, q = 0;
                int32_t mv = 0;
                if (p < total) {
                    r = p / J.nQ;
                    q = p - r * J.nQ;
                    mv = J.matches[r * J.ld + q];
                }
                unsigned long long mask = __ballot(p < total && mv == -1);
                while (mask) {
                    const int f = __builtin_ctzll(mask);
                    const int64_t rf = __shfl(r, f), qf = __shfl(q, f);
                    giant_pair(J, rf, qf, used_r, used_q, bv, bij);
                    mask &= mask - 1ull;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kWave) k_giant(HeavyArgs a) { giant_body(a, nullptr, 1); }

__global__ void __launch_bounds__(kWave)
k_giant_batch(const HeavyArgs *__restrict__ jobs, int n_jobs) {
    giant_body(jobs[0], jobs, n_jobs);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct Layout {
    // query index (written by PREPARE, read-only afterwards)
    int32_t *tile_np;
    int64_t *tile_pbase;
    float *tile_lo, *tile_iw, *smz;
    uint32_t *swho;
    int32_t *dir;
    float *sint;
    int32_t *tile_flags;
    uint32_t *tile_order;  // tiles by peak count, descending (order_tiles)
    // per-call scratch (LIGHT / HEAVY)
    unsigned long long *counters, *heavy_queue, *giant_queue;
    int64_t ntiles, heavy_cap;
    size_t index_bytes, scratch_bytes;
};

// The query index and the call scratch are carved from two bases.  Their
// sizes depend on the queries only (the heavy queue has a fixed capacity;
// pairs beyond it are found by scanning for the -1 sentinel instead), so one
// prepared index serves any reference set, and calls with their own scratch
// may share it.
Layout layout_n(void *index_base, void *scratch_base, int64_t ntiles, int64_t npeaks);

Layout layout(void *index_base, void *scratch_base, const cs_spectra &Q) {
    return layout_n(index_base, scratch_base, (Q.n_spectra + kTileQueries - 1) / kTileQueries,
                    Q.n_peaks);
}

// ntiles query tiles holding npeaks peaks in all (one job or a whole batch).
Layout layout_n(void *index_base, void *scratch_base, int64_t ntiles, int64_t npeaks) {
    Layout L{};
    L.ntiles = ntiles;
    const cs_spectra Q{0, npeaks, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0.0f};
    L.heavy_cap = kHeavyQueueCap;
    const int64_t T = L.ntiles > 0 ? L.ntiles : 1;
    Carver c(index_base);
    L.tile_np = c.take<int32_t>(T);
    L.tile_pbase = c.take<int64_t>(T);
    L.tile_lo = c.take<float>(T);
    L.tile_iw = c.take<float>(T);
    L.smz = c.take<float>(Q.n_peaks + T);
    L.swho = c.take<uint32_t>(Q.n_peaks + T);
    L.dir = c.take<int32_t>(T * (kTileBuckets + 1));
    L.sint = c.take<float>(Q.n_peaks + T);
    L.tile_flags = c.take<int32_t>(T);
    L.tile_order = c.take<uint32_t>(T);
    L.index_bytes = align_up(c.used, 256);
    Carver d(scratch_base);
    L.counters = d.take<unsigned long long>(kCounters);
    L.heavy_queue = d.take<unsigned long long>(L.heavy_cap);
    L.giant_queue = d.take<unsigned long long>(L.heavy_cap);
    L.scratch_bytes = align_up(d.used, 256);
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

// Dynamic LDS bytes of the lane kernel for light_candidates K and tile
// capacity `cap` (K as the launchers round it: 1, 4, 16, else 8).
size_t light_lds_bytes(int K, int cap) {
    switch (K) {
        case 1: return LightLayout<1>::bytes(cap);
        case 4: return LightLayout<4>::bytes(cap);
        case 16: return LightLayout<16>::bytes(cap);
        default: return LightLayout<8>::bytes(cap);
    }
}
int light_layout_cap(int K) {
    switch (K) {
        case 1: return LightLayout<1>::cap;
        case 4: return LightLayout<4>::cap;
        case 16: return LightLayout<16>::cap;
        default: return LightLayout<8>::cap;
    }
}

constexpr const char *kDeviceQueryFailed =
    "HIP device query failed: no usable device for the HIP runtime this library runs on";

// Tile capacity of k_light for K, 0 when the device offers less LDS per
// workgroup than the kernel's compile-time layout needs, or -1 when the HIP
// device query itself fails.
int light_capacity(int K) {
    static int max_lds = -1;  // cached once a query succeeds
    if (max_lds < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return -1;
        max_lds = v;
    }
    const int cap = light_layout_cap(K);
    return (size_t)max_lds >= light_lds_bytes(K, cap) ? cap : 0;
}

template <int K, bool F32>
cs_status launch_light(dim3 grid, hipStream_t st, LightArgs a) {
    a.lds_cap = light_capacity(K);
    if (a.lds_cap < 0) return fail(CS_ERR_HIP, "%s", kDeviceQueryFailed);
    if (a.lds_cap == 0)
        return fail(CS_ERR_UNSUPPORTED, "device has less than %zu bytes of LDS per workgroup",
                    LightLayout<K>::bytes(LightLayout<K>::cap));
    const size_t lds = LightLayout<K>::bytes(a.lds_cap);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_light<K, F32>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return fail(CS_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
        attr_set = true;
    }
    hipLaunchKernelGGL((k_light<K, F32>), grid, dim3(kLightThreads), lds, st, a);
    CS_LAUNCH_CHECK("k_light");
    return CS_OK;
}

// Compute units of the current device (cached per device; 256 if unknown).
int cu_count() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] > 0) return cache[dev];
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v < 1)
        v = 256;
    cache[dev] = v;
    return v;
}

template <int K, bool F32>
cs_status launch_light_sched(hipStream_t st, LightArgs a, const uint32_t *order, int64_t n_items,
                             unsigned long long *next) {
    a.lds_cap = light_capacity(K);
    if (a.lds_cap < 0) return fail(CS_ERR_HIP, "%s", kDeviceQueryFailed);
    if (a.lds_cap == 0)
        return fail(CS_ERR_UNSUPPORTED, "device has less than %zu bytes of LDS per workgroup",
                    LightLayout<K>::bytes(LightLayout<K>::cap));
    const size_t lds = LightLayout<K>::bytes(a.lds_cap);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_light_sched<K, F32>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return fail(CS_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
        attr_set = true;
    }
    // as many workgroups as the compute units hold at once (the LDS decides)
    const int64_t per_cu = (int64_t)(160 * 1024 / (lds + 1024)) > 0 ? 160 * 1024 / (lds + 1024) : 1;
    const int64_t grid = std::min<int64_t>(n_items, per_cu * cu_count());
    hipLaunchKernelGGL((k_light_sched<K, F32>), dim3((unsigned)grid), dim3(kLightThreads), lds, st,
                       a, order, n_items, next);
    CS_LAUNCH_CHECK("k_light_sched");
    return CS_OK;
}

template <bool F32>
cs_status launch_light_sched_k(int K, hipStream_t st, const LightArgs &a, const uint32_t *order,
                               int64_t n_items, unsigned long long *next) {
    switch (K) {
        case 1: return launch_light_sched<1, F32>(st, a, order, n_items, next);
        case 4: return launch_light_sched<4, F32>(st, a, order, n_items, next);
        case 16: return launch_light_sched<16, F32>(st, a, order, n_items, next);
        default: return launch_light_sched<8, F32>(st, a, order, n_items, next);
    }
}

template <bool F32>
cs_status launch_light_k(int K, dim3 grid, hipStream_t st, const LightArgs &a) {
    switch (K) {
        case 1: return launch_light<1, F32>(grid, st, a);
        case 4: return launch_light<4, F32>(grid, st, a);
        case 16: return launch_light<16, F32>(grid, st, a);
        default: return launch_light<8, F32>(grid, st, a);
    }
}

// References per light workgroup: more amortise the tile staging and the
// workgroup's tail (its waves wait for the slowest), fewer keep enough
// workgroups for the chip when there are few query tiles.
int default_refs_per_block(int64_t nR, int64_t ntiles) {
    for (int rpb = kMaxLightRefs; rpb > 64; rpb >>= 1)
        if (ntiles * ((nR + rpb - 1) / rpb) >= 512) return rpb;
    return 64;
}

// Lane-kernel arguments of one job; `tiles` is the job's first tile in L.
LightArgs make_light(const cs_spectra &Rs, const cs_spectra &Qs, double tol, double shift,
                     int64_t ml, float *score, int32_t *matches, uint8_t *overflow, int64_t ld,
                     const Layout &L, int64_t tiles, int64_t nrb, int rpb) {
    LightArgs la{};
    la.R = view(Rs);
    la.Q = view(Qs);
    la.nR = Rs.n_spectra;
    la.nQ = Qs.n_spectra;
    la.tile_np = L.tile_np + tiles;
    la.tile_pbase = L.tile_pbase + tiles;
    la.tile_lo = L.tile_lo + tiles;
    la.tile_iw = L.tile_iw + tiles;
    la.smz = L.smz;
    la.swho = L.swho;
    la.dir = L.dir + tiles * (kTileBuckets + 1);
    la.sint = L.sint;
    la.tile_flags = L.tile_flags + tiles;
    la.no_mz_pow = (Rs.mz_pow == nullptr && Qs.mz_pow == nullptr) ? 1 : 0;
    const float mzmax = Rs.max_abs_mz > Qs.max_abs_mz ? Rs.max_abs_mz : Qs.max_abs_mz;
    la.tolc = coarse_tolerance(tol, shift, mzmax);
    la.tol = tol;
    la.shift = shift;
    la.exact_from_tile = (Qs.mz_exact == nullptr && shift == 0.0) ? 1 : 0;
    la.match_limit = ml;
    la.score = score;
    la.matches = matches;
    la.overflow = overflow;
    la.ld = ld;
    la.heavy_queue = L.heavy_queue;
    la.heavy_cap = L.heavy_cap;
    la.counters = L.counters;
    la.n_ref_blocks = nrb;
    la.refs_per_block = rpb;
    return la;
}

// float32 path: m/z exact in float32 on both sides, no shift, mz_power 0
bool light_f32(const LightArgs &la, const cs_spectra &Rs) {
    return la.exact_from_tile && Rs.mz_exact == nullptr && la.no_mz_pow;
}

// Cooperative-kernel arguments of one job; r_words / q_words / stage are the
// LDS shape every job of a launch shares.
HeavyArgs make_heavy(const cs_spectra &Rs, const cs_spectra &Qs, double tol, double shift,
                     int64_t ml, float *score, int32_t *matches, uint8_t *overflow, int64_t ld,
                     const Layout &L, int T, int r_words, int q_words, int stage) {
    HeavyArgs a{};
    a.R = view(Rs);
    a.Q = view(Qs);
    a.nR = Rs.n_spectra;
    a.nQ = Qs.n_spectra;
    const float mzmax = Rs.max_abs_mz > Qs.max_abs_mz ? Rs.max_abs_mz : Qs.max_abs_mz;
    a.tolc = coarse_tolerance(tol, shift, mzmax);
    a.shift32 = (float)shift;
    a.tol = tol;
    a.shift = shift;
    a.match_limit = ml;
    a.round_cap = T;
    int cap = 1;
    while (cap < 2 * T + kWave) cap <<= 1;
    a.buf_cap = cap;
    a.r_words = r_words;
    a.q_words = q_words;
    a.stage = stage;
    a.exact_from_stage = (Qs.mz_exact == nullptr && shift == 0.0) ? 1 : 0;
    a.score = score;
    a.matches = matches;
    a.overflow = overflow;
    a.ld = ld;
    a.heavy_queue = L.heavy_queue;
    a.heavy_cap = L.heavy_cap;
    a.counters = L.counters;
    a.giant_queue = L.giant_queue;
    return a;
}

// LDS bytes of k_heavy / k_giant for a shared shape.
size_t heavy_lds_bytes(int r_words, int q_words, int stage) {
    const int S = stage ? kHeavyStage : 0;
    return (size_t)(kHeavyBufPad + 2 * S) * 12 + sizeof(uint32_t) * (size_t)(r_words + q_words);
}
size_t giant_lds_bytes(int buf_cap, int r_words, int q_words) {
    return (size_t)buf_cap * (sizeof(double) + sizeof(uint32_t)) +
           sizeof(uint32_t) * (size_t)(r_words + q_words);
}

cs_status set_lds_attr(const void *fn, size_t bytes, size_t &have) {
    if (bytes <= have) return CS_OK;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail(CS_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
    have = bytes;
    return CS_OK;
}

// The query-tile index (all of L's tiles), for one job or a batch.
cs_status launch_prepare(const PrepJob &one, const PrepJob *jobs, const int32_t *tile_job,
                         const Layout &L, hipStream_t hs, bool order) {
    auto blocks = [](int64_t n, int per) {
        int64_t b = (n + per - 1) / per;
        return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
    };
    hipLaunchKernelGGL(k_tile_meta, dim3(blocks(L.ntiles, 256 / kWave)), dim3(256), 0, hs, one, jobs,
                       tile_job, L.ntiles, L.tile_np, L.tile_pbase, L.tile_lo, L.tile_iw);
    CS_LAUNCH_CHECK("k_tile_meta");
    const size_t ilds = sizeof(int) * (2 * kTileBuckets + 1) + 8 * (size_t)kIndexStage;
    static bool index_attr = false;
    if (!index_attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tile_index),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)ilds);
        if (e != hipSuccess) return fail(CS_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
        index_attr = true;
    }
    hipLaunchKernelGGL(k_tile_index, dim3(blocks(L.ntiles, 1)), dim3(kIndexThreads), ilds, hs, one,
                       jobs, tile_job, L.ntiles, L.tile_np, L.tile_pbase, L.tile_lo, L.tile_iw,
                       L.smz, L.swho, L.dir);
    CS_LAUNCH_CHECK("k_tile_index");
    hipLaunchKernelGGL(k_tile_sort, dim3(blocks(L.ntiles, 1)), dim3(256), 0, hs, one, jobs,
                       tile_job, L.ntiles, L.tile_np, L.tile_pbase, L.tile_lo, L.tile_iw, L.smz,
                       L.swho, L.dir);
    CS_LAUNCH_CHECK("k_tile_sort");
    // one more workgroup writes the tile order (a batch orders its own items)
    hipLaunchKernelGGL(k_tile_values, dim3(blocks(L.ntiles, 1) + (order ? 1 : 0)), dim3(1024), 0,
                       hs, one, jobs, tile_job, L.ntiles, L.tile_np, L.tile_pbase, L.swho, L.sint,
                       L.tile_flags, order ? L.tile_order : nullptr);
    CS_LAUNCH_CHECK("k_tile_values");
    return CS_OK;
}

// One whole-pipeline or staged call with the index and scratch at two bases.
cs_status cosine_greedy_run(const cs_spectra *refs, const cs_spectra *queries,
                            const cs_cosine_params *params, float *score, int32_t *matches,
                            uint8_t *overflow, int64_t ld, const Layout &L, void *stream) {
    cs_status st;
    const cs_spectra &Rs = *refs, &Qs = *queries;
    const int64_t nR = Rs.n_spectra, nQ = Qs.n_spectra;
    const double tol = params->tolerance, shift = params->shift;
    int K = params->light_candidates ? params->light_candidates : kDefaultK;
    const int T = params->round_capacity ? params->round_capacity : kDefaultRound;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const SpecView Q = view(Qs);
    const float shift32 = (float)shift;
    const int stages = params->stages ? params->stages
                                      : (CS_STAGE_PREPARE | CS_STAGE_LIGHT | CS_STAGE_HEAVY);
    if (nQ == 0) return CS_OK;
    if (stages & CS_STAGE_PREPARE) {
        const PrepJob one{Q.mz, Q.off, Q.inp, nQ, shift32, 0, 0};
        cs_status pst = launch_prepare(one, nullptr, nullptr, L, hs, true);
        if (pst != CS_OK) return pst;
    }
    if (nR == 0) return CS_OK;

    const int rpb = params->refs_per_block ? params->refs_per_block : default_refs_per_block(nR, L.ntiles);
    const int64_t nrb = (nR + rpb - 1) / rpb;
    if (nrb * L.ntiles > INT_MAX)
        return fail(CS_ERR_UNSUPPORTED, "grid too large (%lld ref blocks x %lld query tiles)",
                    (long long)nrb, (long long)L.ntiles);
    const int64_t ml = params->match_limit;
    if (stages & CS_STAGE_LIGHT) {
        // the heavy-pair queue of this call starts empty
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, hs, L.counters, (int64_t)kCounters);
        CS_LAUNCH_CHECK("k_fill_u64");
        const LightArgs la = make_light(Rs, Qs, tol, shift, ml, score, matches, overflow, ld, L,
                                        0, nrb, rpb);
        // A few rounds of workgroups per compute unit: the items go longest
        // first from a counter (the last round would otherwise leave compute
        // units idle while a few long items finish).  Many rounds: one
        // workgroup per item, in grid order.
        const int64_t n_items = nrb * L.ntiles;
        const int64_t cus = cu_count();
        const bool f32 = light_f32(la, Rs);
        if (n_items > cus && n_items <= 8 * cus && L.ntiles <= kOrderTiles) {
            st = f32 ? launch_light_sched_k<true>(K, hs, la, L.tile_order, n_items, &L.counters[5])
                     : launch_light_sched_k<false>(K, hs, la, L.tile_order, n_items,
                                                   &L.counters[5]);
        } else {
            const dim3 grid((unsigned)n_items);
            st = f32 ? launch_light_k<true>(K, grid, hs, la) : launch_light_k<false>(K, grid, hs, la);
        }
        if (st != CS_OK) return st;
    }
    if (!(stages & CS_STAGE_HEAVY)) return CS_OK;

    const int r_words = (Rs.max_peaks + 31) / 32 + 1, q_words = (Qs.max_peaks + 31) / 32 + 1;
    const int stage = (Rs.max_peaks <= kHeavyStage && Qs.max_peaks <= kHeavyStage) ? 1 : 0;
    const HeavyArgs a = make_heavy(Rs, Qs, tol, shift, ml, score, matches, overflow, ld, L, T,
                                   r_words, q_words, stage);
    const size_t hlds = heavy_lds_bytes(r_words, q_words, stage);
    if (hlds > 160 * 1024)
        return fail(CS_ERR_UNSUPPORTED, "max_peaks %d/%d needs %zu bytes of LDS", Rs.max_peaks,
                    Qs.max_peaks, hlds);
    static size_t heavy_attr[2] = {0, 0};
    const void *heavy_fn = stage ? reinterpret_cast<const void *>(&k_heavy<true>)
                                 : reinterpret_cast<const void *>(&k_heavy<false>);
    if ((st = set_lds_attr(heavy_fn, hlds, heavy_attr[stage])) != CS_OK) return st;
    if (stage)
        hipLaunchKernelGGL(k_heavy<true>, dim3(kHeavyBlocks), dim3(kHeavyThreads), hlds, hs, a);
    else
        hipLaunchKernelGGL(k_heavy<false>, dim3(kHeavyBlocks), dim3(kHeavyThreads), hlds, hs, a);
    CS_LAUNCH_CHECK("k_heavy");
    const size_t glds = giant_lds_bytes(a.buf_cap, r_words, q_words);
    if (glds > 65536)
        return fail(CS_ERR_UNSUPPORTED,
                    "round_capacity %d with max_peaks %d/%d needs %zu bytes of LDS (> 64 KiB)", T,
                    Rs.max_peaks, Qs.max_peaks, glds);
    hipLaunchKernelGGL(k_giant, dim3(kGiantGrid), dim3(kWave), glds, hs, a);
    CS_LAUNCH_CHECK("k_giant");
    return CS_OK;
}

template <int K, bool F32>
size_t &light_batch_attr() {
    static size_t have = 0;
    return have;
}

// LDS attribute of the batch lane kernel, set once outside any graph capture.
template <int K, bool F32>
cs_status prime_light_batch(int cap) {
    return set_lds_attr(reinterpret_cast<const void *>(&k_light_batch<K, F32>),
                        LightLayout<K>::bytes(cap), light_batch_attr<K, F32>());
}

template <bool F32>
cs_status prime_light_batch_k(int K, int cap) {
    switch (K) {
        case 1: return prime_light_batch<1, F32>(cap);
        case 4: return prime_light_batch<4, F32>(cap);
        case 16: return prime_light_batch<16, F32>(cap);
        default: return prime_light_batch<8, F32>(cap);
    }
}

template <int K, bool F32>
void launch_light_batch(unsigned grid, hipStream_t st, const LightArgs *jobs,
                        const unsigned long long *items, int64_t n, unsigned long long *next,
                        int cap) {
    hipLaunchKernelGGL((k_light_batch<K, F32>), dim3(grid), dim3(kLightThreads),
                       LightLayout<K>::bytes(cap), st, jobs, items, n, next);
}

template <bool F32>
void launch_light_batch_k(int K, unsigned grid, hipStream_t st, const LightArgs *jobs,
                          const unsigned long long *items, int64_t n,
                          unsigned long long *next, int cap) {
    switch (K) {
        case 1: launch_light_batch<1, F32>(grid, st, jobs, items, n, next, cap); break;
        case 4: launch_light_batch<4, F32>(grid, st, jobs, items, n, next, cap); break;
        case 16: launch_light_batch<16, F32>(grid, st, jobs, items, n, next, cap); break;
        default: launch_light_batch<8, F32>(grid, st, jobs, items, n, next, cap); break;
    }
}

// Argument checks shared by the entry points (no HIP call).
cs_status check_call(const cs_spectra *refs, const cs_spectra *queries,
                     const cs_cosine_params *params, const float *score, const int32_t *matches,
                     const uint8_t *overflow, int64_t ld) {
    cs_status st;
    if ((st = check_spectra(refs, "refs")) != CS_OK) return st;
    if ((st = check_spectra(queries, "queries")) != CS_OK) return st;
    if (!params) return fail(CS_ERR_INVALID, "params is NULL");
    const int stages = params->stages;
    if (stages & ~(CS_STAGE_PREPARE | CS_STAGE_LIGHT | CS_STAGE_HEAVY))
        return fail(CS_ERR_INVALID, "unknown stage bits 0x%x", stages);
    const bool computes = stages == 0 || (stages & (CS_STAGE_LIGHT | CS_STAGE_HEAVY));
    if (computes && refs->n_spectra > 0 && queries->n_spectra > 0) {
        if (!score || !matches || !overflow) return fail(CS_ERR_INVALID, "output pointer is NULL");
        if (ld < queries->n_spectra)
            return fail(CS_ERR_INVALID, "ld %lld < number of queries %lld", (long long)ld,
                        (long long)queries->n_spectra);
    }
    const double tol = params->tolerance, shift = params->shift;
    if (!(tol >= 0.0) || tol == __builtin_huge_val())
        return fail(CS_ERR_INVALID, "tolerance must be finite and >= 0");
    if (!(shift == shift) || shift == __builtin_huge_val() || shift == -__builtin_huge_val())
        return fail(CS_ERR_INVALID, "shift must be finite");
    const int K = params->light_candidates ? params->light_candidates : kDefaultK;
    if (K != 1 && K != 4 && K != 8 && K != 16)
        return fail(CS_ERR_INVALID, "light_candidates must be 1, 4, 8 or 16 (got %d)", K);
    const int T = params->round_capacity ? params->round_capacity : kDefaultRound;
    if (T < 4 || T > 1024 || (T & (T - 1)))
        return fail(CS_ERR_INVALID, "round_capacity must be a power of two in [4, 1024]");
    if (params->refs_per_block < 0 || params->refs_per_block > kMaxLightRefs)
        return fail(CS_ERR_INVALID, "refs_per_block must be in [0, %d]", kMaxLightRefs);
    if (computes) {
        // the cooperative kernels' LDS, checked before anything is launched:
        // a light stage without its heavy stage would leave -1 sentinels
        const int r_words = (refs->max_peaks + 31) / 32 + 1;
        const int q_words = (queries->max_peaks + 31) / 32 + 1;
        const int stage = (refs->max_peaks <= kHeavyStage && queries->max_peaks <= kHeavyStage);
        int cap = 1;
        while (cap < 2 * T + kWave) cap <<= 1;
        const size_t hlds = heavy_lds_bytes(r_words, q_words, stage);
        const size_t glds = giant_lds_bytes(cap, r_words, q_words);
        if (hlds > 160 * 1024 || glds > 65536)
            return fail(CS_ERR_UNSUPPORTED,
                        "max_peaks %d/%d with round_capacity %d need %zu / %zu bytes of LDS "
                        "(limits 160 KiB / 64 KiB)",
                        refs->max_peaks, queries->max_peaks, T, hlds, glds);
    }
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
    const Layout L = layout(nullptr, nullptr, *queries);
    *bytes = L.index_bytes + L.scratch_bytes;
    return CS_OK;
}

extern "C" cs_status cs_cosine_greedy(const cs_spectra *refs, const cs_spectra *queries,
                                      const cs_cosine_params *params, float *score,
                                      int32_t *matches, uint8_t *overflow, int64_t ld,
                                      void *workspace, size_t workspace_bytes, void *stream) {
    cs_status st;
    if ((st = check_call(refs, queries, params, score, matches, overflow, ld)) != CS_OK) return st;
    if (queries->n_spectra == 0) return CS_OK;
    if (refs->n_spectra == 0 && !(params->stages & CS_STAGE_PREPARE)) return CS_OK;
    const Layout need = layout(nullptr, nullptr, *queries);
    const size_t total = need.index_bytes + need.scratch_bytes;
    if (!workspace || workspace_bytes < total)
        return fail(CS_ERR_WORKSPACE, "workspace of %zu bytes < %zu required", workspace_bytes,
                    total);
    char *base = static_cast<char *>(workspace);
    const Layout L = layout(base, base + need.index_bytes, *queries);
    return cosine_greedy_run(refs, queries, params, score, matches, overflow, ld, L, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the workspace is a file-scope device array
// ---------------------------------------------------------------------------

// The workspace of cs_cosine_greedy_noalloc.  Every GPU the library runs on
// holds its own copy from module load until the process ends.  In host code
// the name is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[CS_NOALLOC_WORKSPACE_BYTES];

namespace cs {
namespace {

constexpr int kMaxDevices = 64;

struct NoAllocSlot {
    std::once_flag once;
    void *ptr = nullptr;
    hipError_t err = hipSuccess;
};
NoAllocSlot g_noalloc[kMaxDevices];

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with that device current) and kept.
cs_status noalloc_workspace(void **out) {
    int dev = -1;
    CS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(CS_ERR_UNSUPPORTED, "device ordinal %d beyond %d", dev, kMaxDevices);
    NoAllocSlot &s = g_noalloc[dev];
    std::call_once(s.once, [&s]() {
        s.err = hipGetSymbolAddress(&s.ptr, HIP_SYMBOL(scratch));
        if (s.err == hipSuccess && s.ptr == nullptr) s.err = hipErrorInvalidSymbol;
    });
    if (s.err != hipSuccess)
        return fail(CS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                    hipGetErrorString(s.err));
    *out = s.ptr;
    return CS_OK;
}

}  // namespace
}  // namespace cs

extern "C" cs_status cs_cosine_greedy_noalloc(const cs_spectra *refs, const cs_spectra *queries,
                                              const cs_cosine_params *params, float *score,
                                              int32_t *matches, uint8_t *overflow, int64_t ld,
                                              void *stream) {
    cs_status st;
    if ((st = check_call(refs, queries, params, score, matches, overflow, ld)) != CS_OK) return st;
    if (queries->n_spectra == 0) return CS_OK;
    if (refs->n_spectra == 0 && !(params->stages & CS_STAGE_PREPARE)) return CS_OK;
    const Layout need = layout(nullptr, nullptr, *queries);
    const size_t total = need.index_bytes + need.scratch_bytes;
    if (total > (size_t)CS_NOALLOC_WORKSPACE_BYTES)
        return fail(CS_ERR_UNSUPPORTED,
                    "%lld queries of %lld peaks need %zu bytes of workspace; the allocation-free "
                    "entry point holds %zu (use cs_cosine_greedy with a workspace)",
                    (long long)queries->n_spectra, (long long)queries->n_peaks, total,
                    (size_t)CS_NOALLOC_WORKSPACE_BYTES);
    void *ws = nullptr;
    if ((st = noalloc_workspace(&ws)) != CS_OK) return st;
    char *base = static_cast<char *>(ws);
    const Layout L = layout(base, base + need.index_bytes, *queries);
    return cosine_greedy_run(refs, queries, params, score, matches, overflow, ld, L, stream);
}

extern "C" cs_status cs_cosine_greedy_indexed_sizes(const cs_spectra *queries, size_t *index_bytes,
                                                    size_t *scratch_bytes) {
    cs_status st;
    if ((st = check_spectra(queries, "queries")) != CS_OK) return st;
    if (!index_bytes || !scratch_bytes) return fail(CS_ERR_INVALID, "size pointer is NULL");
    const Layout L = layout(nullptr, nullptr, *queries);
    *index_bytes = L.index_bytes;
    *scratch_bytes = L.scratch_bytes;
    return CS_OK;
}

extern "C" cs_status cs_cosine_greedy_indexed(const cs_spectra *refs, const cs_spectra *queries,
                                              const cs_cosine_params *params, float *score,
                                              int32_t *matches, uint8_t *overflow, int64_t ld,
                                              void *index, size_t index_bytes, void *scratch,
                                              size_t scratch_bytes, void *stream) {
    cs_status st;
    if ((st = check_call(refs, queries, params, score, matches, overflow, ld)) != CS_OK) return st;
    if (queries->n_spectra == 0) return CS_OK;
    if (refs->n_spectra == 0 && !(params->stages & CS_STAGE_PREPARE)) return CS_OK;
    const Layout need = layout(nullptr, nullptr, *queries);
    if (!index || index_bytes < need.index_bytes)
        return fail(CS_ERR_WORKSPACE, "index of %zu bytes < %zu required", index_bytes,
                    need.index_bytes);
    const int stages = params->stages ? params->stages
                                      : (CS_STAGE_PREPARE | CS_STAGE_LIGHT | CS_STAGE_HEAVY);
    if ((stages & (CS_STAGE_LIGHT | CS_STAGE_HEAVY)) && (!scratch || scratch_bytes < need.scratch_bytes))
        return fail(CS_ERR_WORKSPACE, "scratch of %zu bytes < %zu required", scratch_bytes,
                    need.scratch_bytes);
    const Layout L = layout(index, scratch, *queries);
    return cosine_greedy_run(refs, queries, params, score, matches, overflow, ld, L, stream);
}

// ---------------------------------------------------------------------------
// Batches of independent jobs: one persistent, work-stealing launch per stage
// ---------------------------------------------------------------------------

struct cs_cosine_batch {
    std::vector<cs_cosine_job> jobs;
    int K = kDefaultK, T = kDefaultRound, rpb = 256, cap = 0;
    std::vector<char> job_f32;  // the job can take the float32 lane kernel
    int64_t n_f32_items = 0;    // items[0, n_f32_items) are float32 jobs' items
    unsigned grid = 0;
    int r_words = 1, q_words = 1, stage = 1, buf_cap = 0;
    int64_t ntiles = 0, npeaks = 0;
    std::vector<int64_t> tile_base, pos_base, nrb;
    std::vector<int32_t> tile_job;
    std::vector<unsigned long long> items;
    size_t off_prep = 0, off_tile_job = 0, off_light = 0, off_heavy = 0, off_items = 0,
           off_index = 0, off_scratch = 0, bytes = 0, hlds = 0, glds = 0;
    char *ws = nullptr;  // the uploaded workspace
    Layout L{};
};

extern "C" cs_status cs_cosine_batch_create(const cs_cosine_job *jobs, int32_t n_jobs,
                                            const cs_batch_params *params,
                                            cs_cosine_batch **batch, size_t *workspace_bytes) {
    if (!batch || !workspace_bytes) return fail(CS_ERR_INVALID, "batch/workspace_bytes is NULL");
    *batch = nullptr;
    if (!jobs || n_jobs < 1 || n_jobs > 4095)
        return fail(CS_ERR_INVALID, "n_jobs must be in [1, 4095] (got %d)", n_jobs);
    cs_batch_params bp{};
    if (params) bp = *params;
    cs_cosine_params cp{};
    cp.light_candidates = bp.light_candidates;
    cp.round_capacity = bp.round_capacity;
    cp.refs_per_block = bp.refs_per_block;
    if (bp.grid < 0) return fail(CS_ERR_INVALID, "grid must be >= 0");
    cs_status st;
    for (int j = 0; j < n_jobs; ++j) {
        const cs_cosine_job &J = jobs[j];
        cp.tolerance = J.tolerance;
        cp.shift = J.shift;
        cp.match_limit = J.match_limit;
        if ((st = check_call(&J.refs, &J.queries, &cp, J.score, J.matches, J.overflow, J.ld)) !=
            CS_OK)
        {
            char msg[256];
            snprintf(msg, sizeof msg, "%s", cs_last_error());
            return fail(st, "job %d: %s", j, msg);
        }
        if (J.refs.n_spectra >= (int64_t(1) << 26) || J.queries.n_spectra >= (int64_t(1) << 26))
            return fail(CS_ERR_UNSUPPORTED, "job %d: batches take at most 2^26 spectra a side", j);
    }
    cs_cosine_batch *b = new cs_cosine_batch();
    b->jobs.assign(jobs, jobs + n_jobs);
    b->K = bp.light_candidates ? bp.light_candidates : kDefaultK;
    b->T = bp.round_capacity ? bp.round_capacity : kDefaultRound;
    b->rpb = bp.refs_per_block ? bp.refs_per_block : 256;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 256;
    b->grid = bp.grid ? (unsigned)bp.grid : (unsigned)cus;
    int maxr = 0, maxq = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const cs_cosine_job &J = jobs[j];
        const bool f = J.queries.mz_exact == nullptr && J.shift == 0.0 &&
                       J.refs.mz_exact == nullptr && J.refs.mz_pow == nullptr &&
                       J.queries.mz_pow == nullptr;
        b->job_f32.push_back(f ? 1 : 0);
        maxr = J.refs.max_peaks > maxr ? J.refs.max_peaks : maxr;
        maxq = J.queries.max_peaks > maxq ? J.queries.max_peaks : maxq;
        const int64_t nt = (J.queries.n_spectra + kTileQueries - 1) / kTileQueries;
        b->tile_base.push_back(b->ntiles);
        b->pos_base.push_back(b->npeaks + b->ntiles);
        for (int64_t t = 0; t < nt; ++t) b->tile_job.push_back(j);
        b->ntiles += nt;
        b->npeaks += J.queries.n_peaks;
        const int64_t nr = J.refs.n_spectra;
        b->nrb.push_back(nr && nt ? (nr + b->rpb - 1) / b->rpb : 0);
        if (b->nrb.back() >= (int64_t(1) << 24)) {
            // a work item carries the reference block in 24 bits
            const long long blocks = (long long)b->nrb.back();
            const int rpb = b->rpb;
            delete b;
            return fail(CS_ERR_UNSUPPORTED,
                        "job %d: %lld reference blocks of %d references (at most 2^24 - 1 per "
                        "job; raise refs_per_block)", j, blocks, rpb);
        }
    }
    if (b->ntiles >= (int64_t(1) << 24) * 1) {
        // tile indices of one job travel in 24 bits; the sum bounds them all
        bool ok = true;
        for (int j = 0; j < n_jobs; ++j)
            ok = ok && (jobs[j].queries.n_spectra + kTileQueries - 1) / kTileQueries < (1 << 24);
        if (!ok) {
            delete b;
            return fail(CS_ERR_UNSUPPORTED, "a job has more than 2^24 query tiles");
        }
    }
    b->r_words = (maxr + 31) / 32 + 1;
    b->q_words = (maxq + 31) / 32 + 1;
    b->stage = (maxr <= kHeavyStage && maxq <= kHeavyStage) ? 1 : 0;
    int cap = 1;
    while (cap < 2 * b->T + kWave) cap <<= 1;
    b->buf_cap = cap;
    b->hlds = heavy_lds_bytes(b->r_words, b->q_words, b->stage);
    b->glds = giant_lds_bytes(b->buf_cap, b->r_words, b->q_words);
    if (b->hlds > 160 * 1024 || b->glds > 65536) {
        delete b;
        return fail(CS_ERR_UNSUPPORTED, "max_peaks %d/%d with round_capacity %d exceed LDS", maxr,
                    maxq, cp.round_capacity);
    }
    // work items (job, reference block, tile), costliest first: references in
    // the block x (their mean peaks + one settle per query) x tile share.  The
    // float32 jobs' items come first: each kernel variant takes its own jobs
    // (one launch per variant), so a job with float64 m/z, a shift or an m/z
    // power does not move the others off the float32 variant.
    std::vector<std::pair<double, unsigned long long>> w;
    for (int j = 0; j < n_jobs; ++j) {
        const cs_cosine_job &J = jobs[j];
        const int64_t nt = (J.queries.n_spectra + kTileQueries - 1) / kTileQueries;
        const double mr = J.refs.n_spectra ? (double)J.refs.n_peaks / J.refs.n_spectra : 0.0;
        for (int64_t rb = 0; rb < b->nrb[j]; ++rb) {
            const int64_t r0 = rb * b->rpb;
            const int64_t nref = r0 + b->rpb < J.refs.n_spectra ? b->rpb : J.refs.n_spectra - r0;
            for (int64_t t = 0; t < nt; ++t) {
                const int64_t q0 = t * kTileQueries;
                const int64_t nq = q0 + kTileQueries < J.queries.n_spectra ? kTileQueries
                                                                          : J.queries.n_spectra - q0;
                const double cost = (double)nref * (mr + (double)nq) * (1.0 + 10.0 * J.tolerance);
                w.emplace_back(cost, ((unsigned long long)j << 48) |
                                         ((unsigned long long)rb << 24) | (unsigned long long)t);
            }
        }
    }
    const std::vector<char> &jf = b->job_f32;
    std::stable_sort(w.begin(), w.end(),
                     [&jf](const std::pair<double, unsigned long long> &x,
                           const std::pair<double, unsigned long long> &y) {
                         const int fx = jf[x.second >> 48], fy = jf[y.second >> 48];
                         return fx != fy ? fx > fy : x.first > y.first;
                     });
    b->items.reserve(w.size());
    for (auto &x : w) {
        b->items.push_back(x.second);
        b->n_f32_items += jf[x.second >> 48];
    }
    // workspace: tables, then the index and the scratch of all jobs together
    Carver c(nullptr);
    c.take<PrepJob>(n_jobs);
    b->off_tile_job = c.used = align_up(c.used, 256);
    c.take<int32_t>(b->ntiles + 1);
    b->off_light = c.used = align_up(c.used, 256);
    c.take<LightArgs>(n_jobs);
    b->off_heavy = c.used = align_up(c.used, 256);
    c.take<HeavyArgs>(n_jobs);
    b->off_items = c.used = align_up(c.used, 256);
    c.take<unsigned long long>(b->items.size() + 1);
    b->off_index = align_up(c.used, 256);
    const Layout need = layout_n(nullptr, nullptr, b->ntiles, b->npeaks);
    b->off_scratch = b->off_index + need.index_bytes;
    b->bytes = b->off_scratch + need.scratch_bytes;
    // LDS attributes now, so that a run can be captured into a graph
    b->cap = light_capacity(b->K);
    if (b->cap < 0) {
        delete b;
        return fail(CS_ERR_HIP, "%s", kDeviceQueryFailed);
    }
    if (b->cap == 0) {
        delete b;
        return fail(CS_ERR_UNSUPPORTED, "device has less than %zu bytes of LDS per workgroup",
                    light_lds_bytes(b->K, light_layout_cap(b->K)));
    }
    static size_t heavy_attr_b[2] = {0, 0}, giant_attr_b = 0;
    const void *heavy_fn = b->stage ? reinterpret_cast<const void *>(&k_heavy_batch<true>)
                                    : reinterpret_cast<const void *>(&k_heavy_batch<false>);
    const int64_t n_items_all = (int64_t)b->items.size();
    if ((b->n_f32_items > 0 && (st = prime_light_batch_k<true>(b->K, b->cap)) != CS_OK) ||
        (b->n_f32_items < n_items_all && (st = prime_light_batch_k<false>(b->K, b->cap)) != CS_OK) ||
        (st = set_lds_attr(heavy_fn, b->hlds, heavy_attr_b[b->stage ? 1 : 0])) != CS_OK ||
        (st = set_lds_attr(reinterpret_cast<const void *>(&k_giant_batch), b->glds,
                           giant_attr_b)) != CS_OK) {
        delete b;
        return st;
    }
    *batch = b;
    *workspace_bytes = b->bytes;
    return CS_OK;
}

extern "C" cs_status cs_cosine_batch_upload(cs_cosine_batch *b, void *workspace,
                                            size_t workspace_bytes, void *stream) {
    if (!b) return fail(CS_ERR_INVALID, "batch is NULL");
    if (!workspace || workspace_bytes < b->bytes)
        return fail(CS_ERR_WORKSPACE, "workspace of %zu bytes < %zu required", workspace_bytes,
                    b->bytes);
    char *ws = static_cast<char *>(workspace);
    b->L = layout_n(ws + b->off_index, ws + b->off_scratch, b->ntiles, b->npeaks);
    const int n = (int)b->jobs.size();
    std::vector<char> host(b->off_index, 0);
    PrepJob *prep = reinterpret_cast<PrepJob *>(host.data());
    int32_t *tile_job = reinterpret_cast<int32_t *>(host.data() + b->off_tile_job);
    LightArgs *light = reinterpret_cast<LightArgs *>(host.data() + b->off_light);
    HeavyArgs *heavy = reinterpret_cast<HeavyArgs *>(host.data() + b->off_heavy);
    unsigned long long *items = reinterpret_cast<unsigned long long *>(host.data() + b->off_items);
    for (int j = 0; j < n; ++j) {
        const cs_cosine_job &J = b->jobs[j];
        const SpecView Q = view(J.queries);
        prep[j] = PrepJob{Q.mz, Q.off, Q.inp, J.queries.n_spectra, (float)J.shift,
                          b->tile_base[j], b->pos_base[j]};
        LightArgs la = make_light(J.refs, J.queries, J.tolerance, J.shift, J.match_limit, J.score,
                                  J.matches, J.overflow, J.ld, b->L, b->tile_base[j], b->nrb[j],
                                  b->rpb);
        la.lds_cap = b->cap;
        la.job = j;
        la.batched = 1;
        light[j] = la;
        HeavyArgs ha = make_heavy(J.refs, J.queries, J.tolerance, J.shift, J.match_limit, J.score,
                                  J.matches, J.overflow, J.ld, b->L, b->T, b->r_words, b->q_words,
                                  b->stage);
        ha.job = j;
        ha.batched = 1;
        heavy[j] = ha;
    }
    for (size_t t = 0; t < b->tile_job.size(); ++t) tile_job[t] = b->tile_job[t];
    for (size_t k = 0; k < b->items.size(); ++k) items[k] = b->items[k];
    hipStream_t hs = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) return fail(CS_ERR_HIP, "batch upload: %s", hipGetErrorString(e));
    b->ws = ws;
    return CS_OK;
}

extern "C" cs_status cs_cosine_batch_run(const cs_cosine_batch *b, void *stream) {
    if (!b) return fail(CS_ERR_INVALID, "batch is NULL");
    if (!b->ws) return fail(CS_ERR_INVALID, "batch is not uploaded");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const Layout &L = b->L;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, hs, L.counters, (int64_t)kCounters);
    CS_LAUNCH_CHECK("k_fill_u64");
    const PrepJob *prep = reinterpret_cast<const PrepJob *>(b->ws);
    const int32_t *tile_job = reinterpret_cast<const int32_t *>(b->ws + b->off_tile_job);
    const LightArgs *light = reinterpret_cast<const LightArgs *>(b->ws + b->off_light);
    const HeavyArgs *heavy = reinterpret_cast<const HeavyArgs *>(b->ws + b->off_heavy);
    const unsigned long long *items =
        reinterpret_cast<const unsigned long long *>(b->ws + b->off_items);
    if (b->ntiles > 0) {
        cs_status st = launch_prepare(PrepJob{}, prep, tile_job, L, hs, false);
        if (st != CS_OK) return st;
    }
    const int64_t n_items = (int64_t)b->items.size();
    if (n_items == 0) return CS_OK;
    // one persistent launch per lane-kernel variant, each over its own jobs'
    // items with its own work counter
    const int64_t nf = b->n_f32_items, nd = n_items - nf;
    if (nf > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>(b->grid, nf);
        launch_light_batch_k<true>(b->K, grid, hs, light, items, nf, &L.counters[5], b->cap);
        CS_LAUNCH_CHECK("k_light_batch");
    }
    if (nd > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>(b->grid, nd);
        launch_light_batch_k<false>(b->K, grid, hs, light, items + nf, nd, &L.counters[6], b->cap);
        CS_LAUNCH_CHECK("k_light_batch");
    }
    const int n = (int)b->jobs.size();
    if (b->stage)
        hipLaunchKernelGGL(k_heavy_batch<true>, dim3(kHeavyBlocks), dim3(kHeavyThreads), b->hlds,
                           hs, heavy, n);
    else
        hipLaunchKernelGGL(k_heavy_batch<false>, dim3(kHeavyBlocks), dim3(kHeavyThreads), b->hlds,
                           hs, heavy, n);
    CS_LAUNCH_CHECK("k_heavy_batch");
    hipLaunchKernelGGL(k_giant_batch, dim3(kGiantGrid), dim3(kWave), b->glds, hs, heavy, n);
    CS_LAUNCH_CHECK("k_giant_batch");
    return CS_OK;
}

extern "C" void cs_cosine_batch_destroy(cs_cosine_batch *b) { delete b; }

