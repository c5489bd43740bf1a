"""The light kernel's per-(reference, tile) logic (csrc/cg_tile.h probe over
the tile peak index + csrc/cg_lane.h candidate slots and register greedy), run
on the CPU through tests/native/tile_harness.cpp, against the oracle.

This is the exact probe, coarse window, float64 exact test, (v desc, i desc,
j desc) insertion and greedy a GPU wave runs; only the sort that builds the
tile index differs (std::sort here, a rank merge on the device, which the GPU
parity tests cover)."""
import ctypes

import numpy as np
import pytest

from cudams.spectra import PackedSpectra


def f32_exact(refs, ques, shift, mzp):
    """The F32 lane kernel's case: float32 m/z on both sides, no shift, mz_power 0."""
    return (shift == 0.0 and mzp == 0.0
            and all(np.array_equal(x.mz.astype(np.float32).astype(np.float64), x.mz)
                    for x in (refs, ques)))


def tile_run(L, K, refs, ques, tol, shift=0.0, mzp=0.0, inp=1.0):
    R, Q = refs.n_spectra, ques.n_spectra
    mx = float(max(np.abs(refs.mz).max(initial=0), np.abs(ques.mz).max(initial=0)))
    score = np.zeros((R, Q), np.float64)
    m = np.zeros((R, Q), np.int32)
    c = np.zeros((R, Q), np.int32)
    h = np.zeros((R, Q), np.int32)
    arr = [np.ascontiguousarray(x) for x in (refs.mz, refs.intensities, refs.offsets, ques.mz,
                                             ques.intensities, ques.offsets)]
    rc = L.tile_scan(K, arr[0].ctypes.data, arr[1].ctypes.data, arr[2].ctypes.data, R,
                     arr[3].ctypes.data, arr[4].ctypes.data, arr[5].ctypes.data, Q, tol, shift,
                     mzp, inp, mx, int(f32_exact(refs, ques, shift, mzp)), score.ctypes.data, m.ctypes.data, c.ctypes.data, h.ctypes.data)
    assert rc == 0
    return score, m, c, h.astype(bool)


def check(L, oracle, K, refs, ques, tol, shift=0.0, mzp=0.0, inp=1.0):
    s, m, c, h = tile_run(L, K, refs, ques, tol, shift, mzp, inp)
    want = oracle.cosine_greedy_matrix((refs.mz, refs.intensities, refs.offsets),
                                       (ques.mz, ques.intensities, ques.offsets), tolerance=tol,
                                       shift=shift, mz_power=mzp, intensity_power=inp)
    wc = want["n_candidates"]
    np.testing.assert_array_equal(c, wc)          # every exact candidate found, no extra
    np.testing.assert_array_equal(h, wc > 64)
    light = ~h
    np.testing.assert_array_equal(m[light], want["matches"][light])
    np.testing.assert_allclose(s[light], want["score"][light], rtol=1e-12, atol=1e-300)
    return int(h.sum())


def grid_spectra(rng, n, n_max=9):
    mzs, its = [], []
    for _ in range(n):
        k = int(rng.integers(0, n_max + 1))
        mzs.append(np.sort(100 + 0.25 * rng.integers(0, 16, k)))
        its.append(rng.integers(1, 4, k) / 4.0)
    return PackedSpectra.from_arrays(mzs, its)


@pytest.mark.parametrize("K", [1, 4, 8, 16])
def test_ties_and_boundaries(tile_harness, oracle, K):
    """Grid m/z (many pairs exactly at +-tol) and 3-valued intensities."""
    rng = np.random.default_rng(K)
    for tol in (0.25, 0.5, 0.1):
        check(tile_harness, oracle, K, grid_spectra(rng, 23), grid_spectra(rng, 70), tol)


@pytest.mark.parametrize("K", [4, 8])
def test_float64_mz_near_window_edge(tile_harness, oracle, K):
    """m/z not representable in float32, within a few ulp of +-tol."""
    rng = np.random.default_rng(100 + K)
    for tol in (0.01, 0.1, 0.0005):
        mzs_r, its_r, mzs_q, its_q = [], [], [], []
        for _ in range(40):
            base = np.sort(rng.uniform(50, 2000, 4))
            eps = rng.integers(-4, 5, 4) * np.spacing(base)
            mzs_r.append(base)
            its_r.append(rng.random(4) + 0.01)
            mzs_q.append(np.sort(base + rng.choice([-1, 1], 4) * tol + eps))
            its_q.append(rng.random(4) + 0.01)
        check(tile_harness, oracle, K, PackedSpectra.from_arrays(mzs_r, its_r),
              PackedSpectra.from_arrays(mzs_q, its_q), tol)


def test_powers_and_shift(tile_harness, oracle):
    rng = np.random.default_rng(7)
    mzs_r = [np.sort(rng.uniform(100, 400, int(k))) for k in rng.integers(1, 12, 30)]
    mzs_q = [np.sort(rng.uniform(100, 400, int(k))) for k in rng.integers(1, 12, 45)]
    refs = PackedSpectra.from_arrays(mzs_r, [rng.random(len(x)) for x in mzs_r])
    for mzp, inp, shift in ((0.5, 2.0, 0.0), (1.0, 0.5, 0.0), (0.0, 1.0, 17.25)):
        ques = PackedSpectra.from_arrays([x - shift for x in mzs_q],
                                         [rng.random(len(x)) for x in mzs_q])
        check(tile_harness, oracle, 16, refs, ques, 3.0, shift, mzp, inp)


def test_uniform_workload_slice(tile_harness, oracle):
    from cudams import synthetic

    heavy = check(tile_harness, oracle, 8, synthetic.uniform_spectra(40, 1),
                  synthetic.uniform_spectra(70, 2), 0.1)
    assert heavy < 40 * 70 // 4


def test_clustered_workload(tile_harness, oracle):
    from cudams import synthetic

    p = synthetic.lognormal_spectra(64, 5, cluster_fraction=0.5)
    assert check(tile_harness, oracle, 8, p, p, 0.01) > 0


def test_coarse_tolerance_bounds(tile_harness):
    for tol in (0.0, 1e-6, 0.01, 0.1, 2.0):
        for m in (0.0, 1.0, 1000.0, 1e5):
            t = tile_harness.coarse_tolerance(tol, 0.0, m)
            assert t > tol
            assert t <= tol + 16 * (m + tol) * 2.0 ** -24 + 1e-30 + 4 * np.spacing(np.float32(tol + 1e-30))


def test_f32_window_matches_float64_test(tile_harness):
    """The F32 lane kernel walks [f32_up(mz_r - tol), f32_down(mz_r + tol)]
    and tests nothing else; that must equal the float64 test of matchms
    find_matches (mz_r - tol <= mz_q <= mz_r + tol, bounds rounded in float64)
    for every float32 m/z.  Checked where it can fail: query m/z on and one
    float32 step either side of both float64 bounds (and of their float32
    roundings), tolerances that are and are not float32 numbers, bounds that
    cross zero."""
    rng = np.random.default_rng(11)
    tols = np.concatenate([[0.0, 0.01, 0.1, 0.5, 1.0, 2.0 ** -10, 1e-7, 3.0, 1e-30],
                           rng.uniform(0, 1, 40), 10.0 ** rng.uniform(-8, 0, 40)])
    mr = np.concatenate([np.float32([0.0, 1e-3, 0.005, 0.01, 50.0, 100.25, 999.99, 1000.0,
                                     65504.0, 3e5]),
                         rng.uniform(0, 2000, 200).astype(np.float32),
                         (2.0 ** rng.integers(-3, 12, 50)).astype(np.float32)])
    R, T, Qs = [], [], []
    for t in tols:
        for m in mr:
            for b in (float(m) - t, float(m) + t):
                c = np.float32(b)
                for base in (c, np.nextafter(c, np.float32(-np.inf)), np.nextafter(c, np.float32(np.inf))):
                    for q in (np.nextafter(base, np.float32(-np.inf)), base,
                              np.nextafter(base, np.float32(np.inf))):
                        R.append(m)
                        T.append(t)
                        Qs.append(q)
    mr_a = np.array(R, np.float32)
    mq_a = np.array(Qs, np.float32)
    tol_a = np.array(T, np.float64)
    # the float64 test must see both outcomes often, or the check pins nothing
    inside = (mq_a.astype(np.float64) <= mr_a.astype(np.float64) + tol_a) & (
        mq_a.astype(np.float64) >= mr_a.astype(np.float64) - tol_a)
    assert 0.2 < inside.mean() < 0.8
    bad = tile_harness.f32_window_mismatches(mr_a.ctypes.data, mq_a.ctypes.data, tol_a.ctypes.data,
                                             len(mr_a))
    assert bad == 0


def test_f32_window_edges_are_exercised(tile_harness, oracle):
    """The grid workload has peaks exactly on both window edges: the float32
    window finds every oracle candidate there, and one a float32 step narrower
    would lose some, so the grid check is sensitive to an off-by-one bound."""
    rng = np.random.default_rng(3)
    refs, ques = grid_spectra(rng, 23), grid_spectra(rng, 70)
    assert f32_exact(refs, ques, 0.0, 0.0)
    want = oracle.cosine_greedy_matrix((refs.mz, refs.intensities, refs.offsets),
                                       (ques.mz, ques.intensities, ques.offsets), tolerance=0.25)
    # tolerance 0.25 on a 0.25 grid: edge peaks are candidates
    _, _, c, _ = tile_run(tile_harness, 8, refs, ques, 0.25)
    np.testing.assert_array_equal(c, want["n_candidates"])
    # the same walk with a tolerance one float32 step short of 0.25 misses them
    _, _, c2, _ = tile_run(tile_harness, 8, refs, ques, float(np.nextafter(np.float32(0.25), 0)))
    assert (c2 < want["n_candidates"]).any()
