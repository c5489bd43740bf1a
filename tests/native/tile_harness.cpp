// Host harness for cosine-similarity_amd/csrc/cg_tile.h + cg_lane.h: builds
// the peak index of each 32-query tile (sorted by (value, lane, peak), bucket
// directory) and runs, per (reference, tile), the probe / exact test /
// per-query candidate slots / register greedy that k_light runs, so
// tests/test_tile_logic.py can compare it with the oracle without a GPU.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "cg_tile.h"

namespace {

constexpr int kMid = 64;  // candidates the device settles per query without the heavy kernel

struct Entry {
    float x;
    int lane, j;
};

template <int K>
void run(const double *rmz, const double *rint, const int64_t *roff, int R, const double *qmz,
         const double *qint, const int64_t *qoff, int Q, double tol, double shift, double mzp,
         double inp, double max_abs_mz, double *score, int *matches, int *candidates,
         int *heavy) {
    const float tolc = cs::coarse_tolerance(tol, shift, max_abs_mz);
    const float shift32 = (float)shift;
    auto wmz = [&](double m) { return mzp == 0.0 ? 1.0 : pow(m, mzp); };
    auto norm = [&](const double *mz, const double *in, int64_t a, int64_t b) {
        double s = 0.0;
        for (int64_t k = a; k < b; ++k) {
            const double w = wmz(mz[k]) * pow(in[k], inp);
            s += w * w;
        }
        return sqrt(s);
    };
    for (int q0 = 0; q0 < Q; q0 += cs::kTileQueries) {
        const int nl = std::min(cs::kTileQueries, Q - q0);
        std::vector<Entry> e;
        for (int l = 0; l < nl; ++l)
            for (int64_t g = qoff[q0 + l]; g < qoff[q0 + l + 1]; ++g)
                e.push_back({(float)qmz[g] + shift32, l, (int)(g - qoff[q0 + l])});
        std::sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) {
            return a.x < b.x || (a.x == b.x && (a.lane < b.lane || (a.lane == b.lane && a.j < b.j)));
        });
        const int np = (int)e.size();
        const float lo = np ? e.front().x : 0.0f, hi = np ? e.back().x : 0.0f;
        const float iw = cs::tile_inv_width(lo, hi);
        std::vector<float> smz(np + 1);
        std::vector<uint32_t> who(np + 1, 0u);
        for (int p = 0; p < np; ++p) {
            smz[p] = e[p].x;
            who[p] = cs::tile_who(e[p].lane, e[p].j);
        }
        smz[np] = __builtin_huge_valf();
        std::vector<int> dir(cs::kTileBuckets + 1);
        for (int p = 0; p <= np; ++p) {
            const int bp = p < np ? cs::tile_bucket(smz[p], lo, iw) : cs::kTileBuckets;
            const int bq = p > 0 ? cs::tile_bucket(smz[p - 1], lo, iw) : -1;
            for (int b = bq + 1; b <= bp; ++b) dir[b] = p;
        }
        for (int r = 0; r < R; ++r) {
            std::vector<int> cnt(nl, 0);
            std::vector<double> sv(nl * kMid);
            std::vector<uint32_t> sij(nl * kMid);
            for (int64_t g = roff[r]; g < roff[r + 1]; ++g) {
                const int i = (int)(g - roff[r]);
                const float ri = (float)rmz[g];
                cs::tile_probe(ri - tolc, ri + tolc, lo, iw, [&](int p) { return smz[p]; },
                               [&](int b) { return dir[b]; }, [&](int p) {
                                   const int ql = (int)(who[p] >> 16), j = (int)(who[p] & 0xffffu);
                                   const int64_t qp = qoff[q0 + ql] + j;
                                   if (!cs::exact_window(rmz[g], qmz[qp], tol, shift)) return;
                                   double v = wmz(rmz[g]) * pow(rint[g], inp);
                                   v = v * wmz(qmz[qp]);
                                   v = v * pow(qint[qp], inp);
                                   const int pos = cnt[ql]++;
                                   if (pos < kMid) {
                                       sv[ql * kMid + pos] = v;
                                       sij[ql * kMid + pos] = cs::pack_ij(i, j);
                                   }
                               });
            }
            for (int l = 0; l < nl; ++l) {
                const int64_t o = (int64_t)r * Q + q0 + l;
                const int n = cnt[l];
                candidates[o] = n;
                heavy[o] = n > kMid;
                if (n > kMid) {
                    score[o] = 0.0;
                    matches[o] = -1;
                    continue;
                }
                double sc = 0.0;
                int m = 0;
                if (n <= K) {  // the lane path: register insertion + greedy
                    cs::LaneCandidates<K> c;
                    c.reset();
                    for (int s = 0; s < n; ++s) c.insert(sv[l * kMid + s], sij[l * kMid + s]);
                    c.greedy(sc, m);
                } else {       // the wave path: sort by key, then sequential greedy
                    std::vector<std::pair<double, uint32_t>> e;
                    for (int s = 0; s < n; ++s) e.push_back({sv[l * kMid + s], sij[l * kMid + s]});
                    std::sort(e.begin(), e.end(), [](const auto &a, const auto &b) {
                        return cs::key_greater(a.first, a.second, b.first, b.second);
                    });
                    std::vector<uint32_t> ui, uj;
                    for (auto &x : e) {
                        const uint32_t a = x.second >> 16, b = x.second & 0xffffu;
                        if (std::find(ui.begin(), ui.end(), a) != ui.end() ||
                            std::find(uj.begin(), uj.end(), b) != uj.end())
                            continue;
                        ui.push_back(a);
                        uj.push_back(b);
                        sc += x.first;
                        ++m;
                    }
                }
                score[o] = n ? sc / (norm(rmz, rint, roff[r], roff[r + 1]) *
                                     norm(qmz, qint, qoff[q0 + l], qoff[q0 + l + 1]))
                             : 0.0;
                matches[o] = m;
            }
        }
    }
}

}  // namespace

extern "C" int tile_scan(int K, const double *rmz, const double *rint, const int64_t *roff, int R,
                         const double *qmz, const double *qint, const int64_t *qoff, int Q,
                         double tol, double shift, double mzp, double inp, double max_abs_mz,
                         double *score, int *matches, int *candidates, int *heavy) {
    switch (K) {
        case 1: run<1>(rmz, rint, roff, R, qmz, qint, qoff, Q, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy); return 0;
        case 4: run<4>(rmz, rint, roff, R, qmz, qint, qoff, Q, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy); return 0;
        case 8: run<8>(rmz, rint, roff, R, qmz, qint, qoff, Q, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy); return 0;
        case 16: run<16>(rmz, rint, roff, R, qmz, qint, qoff, Q, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy); return 0;
        default: return -1;
    }
}

extern "C" float coarse_tolerance(double tol, double shift, double max_abs_mz) {
    return cs::coarse_tolerance(tol, shift, max_abs_mz);
}
