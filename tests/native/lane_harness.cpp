// Host harness for cosine-similarity_amd/csrc/cg_lane.h: runs the per-lane
// scan the light kernel runs, on the CPU, with the same float32 rounding of
// m/z, the same coarse tolerance and the same float64 exact test, so
// tests/test_lane_logic.py can compare it with the oracle without a GPU.
#include <math.h>
#include <stdint.h>
#include <vector>

#include "cg_lane.h"

namespace {

template <int K>
int run(const double *rmz, const double *rint, int nr, const double *qmz, const double *qint,
        int nq, double tol, double shift, double mzp, double inp, double max_abs_mz,
        double *score, int *matches, int *candidates, int *heavy) {
    std::vector<float> r32(nr), q32(nq + 1);
    const float shift32 = (float)shift;
    for (int k = 0; k < nr; ++k) r32[k] = (float)rmz[k];
    for (int j = 0; j < nq; ++j) q32[j] = (float)qmz[j] + shift32;
    q32[nq] = __builtin_huge_valf();
    auto wmz = [&](double m) { return mzp == 0.0 ? 1.0 : pow(m, mzp); };
    const float tolc = cs::coarse_tolerance(tol, shift, max_abs_mz);
    cs::LaneCandidates<K> cand;
    auto ref_mz = [&](int k) { return r32[k]; };
    auto qry_mz = [&](int j) { return q32[j]; };
    auto exact = [&](int k, int j) { return cs::exact_window(rmz[k], qmz[j], tol, shift); };
    auto value = [&](int k, int j) {
        double v = wmz(rmz[k]) * pow(rint[k], inp);
        v = v * wmz(qmz[j]);
        v = v * pow(qint[j], inp);
        return v;
    };
    cs::LaneResult res = cs::scan_pair<K>(nr, nq, tolc, ref_mz, qry_mz, exact, value, cand);
    double sr = 0.0, sq = 0.0;
    for (int k = 0; k < nr; ++k) { double w = wmz(rmz[k]) * pow(rint[k], inp); sr += w * w; }
    for (int j = 0; j < nq; ++j) { double w = wmz(qmz[j]) * pow(qint[j], inp); sq += w * w; }
    *score = res.candidates && !res.heavy ? res.score / (sqrt(sr) * sqrt(sq)) : 0.0;
    *matches = res.matches;
    *candidates = res.candidates;
    *heavy = res.heavy ? 1 : 0;
    return 0;
}

}  // namespace

extern "C" int lane_scan(int K, const double *rmz, const double *rint, int nr, const double *qmz,
                         const double *qint, int nq, double tol, double shift, double mzp,
                         double inp, double max_abs_mz, double *score, int *matches,
                         int *candidates, int *heavy) {
    switch (K) {
        case 1: return run<1>(rmz, rint, nr, qmz, qint, nq, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy);
        case 4: return run<4>(rmz, rint, nr, qmz, qint, nq, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy);
        case 8: return run<8>(rmz, rint, nr, qmz, qint, nq, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy);
        case 16: return run<16>(rmz, rint, nr, qmz, qint, nq, tol, shift, mzp, inp, max_abs_mz, score, matches, candidates, heavy);
        default: return -1;
    }
}

extern "C" float coarse_tolerance(double tol, double shift, double max_abs_mz) {
    return cs::coarse_tolerance(tol, shift, max_abs_mz);
}
