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
                     mzp, inp, mx, score.ctypes.data, m.ctypes.data, c.ctypes.data, h.ctypes.data)
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
