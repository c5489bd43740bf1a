"""The per-lane scan of the light kernel (csrc/cg_lane.h), run on the CPU
through tests/native/lane_harness.cpp, against the oracle.

This exercises the exact code a GPU lane runs: float32 merge with the coarse
tolerance, float64 exact test, register insertion sort with the
(v desc, i desc, j desc) key and the register greedy.
"""
import ctypes

import numpy as np
import pytest


def lane(L, K, mz1, i1, mz2, i2, tol, shift=0.0, mzp=0.0, inp=1.0):
    a, b, c, d = (np.ascontiguousarray(x, np.float64) for x in (mz1, i1, mz2, i2))
    mx = float(max(np.abs(a).max(initial=0), np.abs(c).max(initial=0)))
    s = ctypes.c_double()
    m, n, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = L.lane_scan(K, a.ctypes.data, b.ctypes.data, len(a), c.ctypes.data, d.ctypes.data,
                     len(c), tol, shift, mzp, inp, mx, ctypes.byref(s), ctypes.byref(m),
                     ctypes.byref(n), ctypes.byref(h))
    assert rc == 0
    return s.value, m.value, n.value, bool(h.value)


def check(L, oracle, K, mz1, i1, mz2, i2, tol, shift=0.0, mzp=0.0, inp=1.0):
    s, m, n, heavy = lane(L, K, mz1, i1, mz2, i2, tol, shift, mzp, inp)
    os_, om, on = oracle.cosine_greedy_pair(mz1, i1, mz2, i2, tolerance=tol, shift=shift,
                                            mz_power=mzp, intensity_power=inp)
    if on > K:
        assert heavy and n == K + 1
        return True
    assert not heavy
    assert n == on and m == om
    assert s == pytest.approx(os_, rel=1e-12, abs=1e-300)
    return False


@pytest.mark.parametrize("K", [1, 4, 8, 16])
def test_lane_vs_oracle_ties_and_boundaries(lane_harness, oracle, K):
    """Grid-valued m/z (many pairs exactly at +-tol) and 3-valued intensities."""
    rng = np.random.default_rng(K)
    heavy = 0
    for _ in range(1500):
        n1, n2 = rng.integers(0, 9, size=2)
        mz1 = np.sort(100 + 0.25 * rng.integers(0, 16, n1))
        mz2 = np.sort(100 + 0.25 * rng.integers(0, 16, n2))
        i1 = rng.integers(1, 4, n1) / 4.0
        i2 = rng.integers(1, 4, n2) / 4.0
        tol = float(rng.choice([0.25, 0.5, 0.1]))
        heavy += check(lane_harness, oracle, K, mz1, i1, mz2, i2, tol)
    assert heavy > 0 or K == 16


@pytest.mark.parametrize("K", [4, 8])
def test_lane_float64_mz_near_window_edge(lane_harness, oracle, K):
    """m/z not representable in float32, placed within a few ulp of +-tol: the
    coarse float32 window must keep every exact float64 candidate."""
    rng = np.random.default_rng(100 + K)
    for _ in range(2000):
        base = rng.uniform(50, 2000, size=4)
        tol = float(rng.choice([0.01, 0.1, 0.0005]))
        eps = rng.integers(-4, 5, size=4) * np.spacing(base)
        mz1 = np.sort(base)
        sign = rng.choice([-1, 1], size=4)
        mz2 = np.sort(mz1 + sign * tol + eps)
        i1 = rng.random(4) + 0.01
        i2 = rng.random(4) + 0.01
        check(lane_harness, oracle, K, mz1, i1, mz2, i2, tol)


def test_lane_powers_and_shift(lane_harness, oracle):
    rng = np.random.default_rng(7)
    for _ in range(500):
        n1, n2 = rng.integers(1, 12, size=2)
        mz1 = np.sort(rng.uniform(100, 400, n1))
        mz2 = np.sort(rng.uniform(100, 400, n2))
        i1 = rng.random(n1)
        i2 = rng.random(n2)
        for mzp, inp, shift in ((0.5, 2.0, 0.0), (1.0, 0.5, 0.0), (0.0, 1.0, 17.25)):
            check(lane_harness, oracle, 16, mz1, i1, mz2 - shift, i2, 3.0, shift, mzp, inp)


def test_lane_uniform_workload_slice(lane_harness, oracle):
    """32 x 32 of the 256 x 256 uniform workload (tol 0.1)."""
    from cudams import synthetic

    R = synthetic.uniform_spectra(32, 1)
    Q = synthetic.uniform_spectra(32, 2)
    heavy = 0
    for r in range(R.n_spectra):
        a = slice(R.offsets[r], R.offsets[r + 1])
        for q in range(Q.n_spectra):
            b = slice(Q.offsets[q], Q.offsets[q + 1])
            heavy += check(lane_harness, oracle, 8, R.mz[a], R.intensities[a], Q.mz[b],
                           Q.intensities[b], 0.1)
    assert heavy < R.n_spectra * Q.n_spectra // 4


def test_coarse_tolerance_bounds(lane_harness):
    for tol in (0.0, 1e-6, 0.01, 0.1, 2.0):
        for m in (0.0, 1.0, 1000.0, 1e5):
            t = lane_harness.coarse_tolerance(tol, 0.0, m)
            assert t > tol
            assert t <= tol + 16 * (m + tol) * 2.0 ** -24 + 1e-30 + 4 * np.spacing(np.float32(tol + 1e-30))
