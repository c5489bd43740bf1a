"""CPUParallelCosineGreedy (the host engine libcsamd_cpu.so) against the
oracle, on the CPU.  Bar: matches bit-exact, score within 1e-5 absolute of the
float64 oracle (in practice the float32 of the same float64 sum).

Reference: cudams/similarity/CPUParallelCosineGreedy.py:19-108 and its numba
kernel spectrum_similarity_functions.py:281-320."""
import ctypes
import json
import os
import warnings

import numpy as np
import pytest

from conftest import GOLDEN

SCORE_ATOL = 1e-5


def as_spectra(p):
    from cudams.spectrum import Spectrum

    return [Spectrum(p.mz[p.offsets[i]:p.offsets[i + 1]],
                     p.intensities[p.offsets[i]:p.offsets[i + 1]]) for i in range(p.n_spectra)]


def oracle_csr(oracle, refs, queries, **kw):
    return oracle.cosine_greedy_matrix((refs.mz, refs.intensities, refs.offsets),
                                       (queries.mz, queries.intensities, queries.offsets), **kw)


def run(refs, queries, **kw):
    from cudams.similarity import CPUParallelCosineGreedy

    return CPUParallelCosineGreedy(**kw).matrix(as_spectra(refs), as_spectra(queries))


def assert_parity(res, want):
    np.testing.assert_array_equal(res["matches"], want["matches"])
    np.testing.assert_allclose(res["score"], want["score"], rtol=0, atol=SCORE_ATOL)


def test_exported_like_the_reference():
    from cudams.similarity import CPUParallelCosineGreedy, __all__

    assert "CPUParallelCosineGreedy" in __all__
    assert CPUParallelCosineGreedy.score_datatype == [("score", np.float32), ("matches", np.int32)]
    cg = CPUParallelCosineGreedy()
    assert (cg.tolerance, cg.mz_power, cg.int_power, cg.batch_size, cg.shift, cg.match_limit) == (
        0.1, 0.0, 1.0, 100, 0, 1024)
    with pytest.warns(UserWarning, match="Batch size larger than 100"):
        CPUParallelCosineGreedy(batch_size=101)
    with pytest.raises(NotImplementedError):
        cg.matrix([], [], array_type="sparse")


@pytest.mark.parametrize("case", range(4))
def test_known_answer_pairs(case):
    """tests/similarity/test_cosine_greedy.py:27-120 of the reference, through pair()."""
    from cudams.similarity import CPUParallelCosineGreedy
    from cudams.spectrum import Spectrum

    with open(os.path.join(GOLDEN, "cosine_greedy_kats.json")) as fh:
        k = json.load(fh)["pairs"][case]
    mz1, i1 = np.array(k["ref_mz"], float), np.array(k["ref_int"], float)
    mz2, i2 = np.array(k["query_mz"], float), np.array(k["query_int"], float)
    cg = CPUParallelCosineGreedy(tolerance=k["tolerance"], mz_power=k["mz_power"],
                                 intensity_power=k["intensity_power"])
    score = cg.pair(Spectrum(mz1, i1), Spectrum(mz2, i2))
    exp = k["expected_matches"]
    p, q = k["mz_power"], k["intensity_power"]
    num = ((mz1[exp[0]] ** p) * (i1[exp[0]] ** q) * (mz2[exp[1]] ** p) * (i2[exp[1]] ** q)).sum()
    den = np.sqrt((((mz1 ** p) * (i1 ** q)) ** 2).sum()) * np.sqrt((((mz2 ** p) * (i2 ** q)) ** 2).sum())
    assert score["score"] == pytest.approx(num / den, 0.0001)
    assert score["matches"] == len(exp[0])


def test_uniform_256(oracle):
    from cudams import synthetic

    refs, ques = synthetic.uniform_spectra(256, 1), synthetic.uniform_spectra(256, 2)
    assert_parity(run(refs, ques, tolerance=0.1), oracle_csr(oracle, refs, ques, tolerance=0.1))


def test_clustered_lognormal(oracle):
    """Clustered spectra: self pairs with thousands of candidates (sort path)."""
    from cudams import synthetic

    p = synthetic.lognormal_spectra(200, 5)
    want = oracle_csr(oracle, p, p, tolerance=0.01)
    assert want["n_candidates"].max() > 1000
    assert_parity(run(p, p, tolerance=0.01), want)


def test_pesticides_against_golden():
    from cudams.similarity import CPUParallelCosineGreedy
    from cudams.spectrum import load_from_mgf, normalize_intensities

    spectra = [normalize_intensities(s)
               for s in load_from_mgf(os.path.join(GOLDEN, "pesticides.mgf"))]
    res = CPUParallelCosineGreedy().matrix(spectra, spectra)
    z = np.load(os.path.join(GOLDEN, "pesticides_oracle.npz"))
    np.testing.assert_array_equal(res["matches"], z["matches"])
    np.testing.assert_allclose(res["score"], z["score"], atol=SCORE_ATOL, rtol=0)


@pytest.mark.parametrize("tol", [0.25, 0.5, 0.0])
def test_window_edges_and_ties_on_a_grid(oracle, tol):
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(17)

    def grid(n):
        mzs, its = [], []
        for _ in range(n):
            k = int(rng.integers(0, 12))
            mzs.append(np.sort(100 + 0.25 * rng.integers(0, 24, k)))
            its.append(rng.integers(1, 4, k) / 4.0)
        return PackedSpectra.from_arrays(mzs, its)

    refs, ques = grid(29), grid(100)
    assert_parity(run(refs, ques, tolerance=tol), oracle_csr(oracle, refs, ques, tolerance=tol))


@pytest.mark.parametrize("mzp,inp,tol", [(0.5, 2.0, 1.0), (1.0, 0.5, 0.1)])
def test_powers(oracle, mzp, inp, tol):
    from cudams import synthetic

    refs, ques = synthetic.uniform_spectra(40, 31), synthetic.uniform_spectra(50, 32)
    assert_parity(run(refs, ques, tolerance=tol, mz_power=mzp, intensity_power=inp),
                  oracle_csr(oracle, refs, ques, tolerance=tol, mz_power=mzp, intensity_power=inp))


def test_float64_mz_and_shift(oracle):
    """m/z that float32 cannot hold, within ulps of the window edge, shifted."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(9)
    mzs_r, its_r, mzs_q, its_q = [], [], [], []
    for _ in range(50):
        base = np.sort(rng.uniform(100, 900, rng.integers(1, 40)))
        mzs_r.append(base)
        its_r.append(rng.random(len(base)) + 0.01)
        eps = rng.integers(-3, 4, len(base)) * np.spacing(base)
        q = np.sort(base + rng.choice([-0.01, 0.01], len(base)) + eps - 3.5)
        mzs_q.append(q)
        its_q.append(rng.random(len(q)) + 0.01)
    refs = PackedSpectra.from_arrays(mzs_r, its_r)
    ques = PackedSpectra.from_arrays(mzs_q, its_q)
    assert_parity(run(refs, ques, tolerance=0.01, shift=3.5),
                  oracle_csr(oracle, refs, ques, tolerance=0.01, shift=3.5))


def test_empty_and_zero_spectra(oracle):
    """Empty spectra score (0, 0); all-zero intensities with a candidate give
    0/0 = NaN, as matchms does."""
    from cudams.spectra import PackedSpectra

    refs = PackedSpectra.from_arrays([[], [100.0, 200.0], [100.0]], [[], [0.5, 1.0], [0.0]])
    ques = PackedSpectra.from_arrays([[100.0], [], [150.0]], [[1.0], [], [0.0]])
    res = run(refs, ques, tolerance=0.1)
    want = oracle_csr(oracle, refs, ques, tolerance=0.1)
    np.testing.assert_array_equal(res["matches"], want["matches"])
    np.testing.assert_array_equal(np.isnan(res["score"]), np.isnan(want["score"]))
    ok = ~np.isnan(want["score"])
    np.testing.assert_allclose(res["score"][ok], want["score"][ok], atol=SCORE_ATOL, rtol=0)
    assert np.isnan(res["score"][2, 0])
    assert run(refs, PackedSpectra.from_arrays([], []), tolerance=0.1).shape == (3, 0)


def _abi_call(refs, queries, tol=0.1):
    from cudams import _native

    L = _native.load_cpu()
    rs, qs = refs.struct(), queries.struct()
    s = np.zeros((refs.n_spectra, max(queries.n_spectra, 1)), np.float32)
    m = np.zeros_like(s, dtype=np.int32)
    return L.cs_cpu_cosine_greedy(ctypes.byref(rs), ctypes.byref(qs), tol, 0.0, s.ctypes.data,
                                  m.ctypes.data, max(queries.n_spectra, 1), 1)


@pytest.mark.parametrize("bad", ["unsorted", "offsets", "max_peaks", "max_abs_mz", "span"])
def test_abi_rejects_corrupt_csr(bad):
    """The C ABI validates host CSR input itself: a caller that bypasses the
    Python packer gets CS_ERR_INVALID, never a silent wrong answer."""
    from cudams import _native
    from cudams.similarity.CPUParallelCosineGreedy import HostSpectra
    from cudams.spectra import PackedSpectra

    good = PackedSpectra.from_arrays([[100.0, 101.0, 102.0], [50.0, 60.0]], [[1, 1, 1], [1, 1]])
    q = HostSpectra(good, 0.0, 1.0)
    H = HostSpectra(good, 0.0, 1.0)
    H.offsets = H.offsets.copy()
    assert _abi_call(H, q) == _native.CS_OK
    if bad == "unsorted":
        H.mz[1], H.mz[2] = H.mz[2], H.mz[1]
    elif bad == "offsets":
        H.offsets[1] = 4
        H.offsets[2] = 3
    elif bad == "max_peaks":
        H.max_peaks = 2
    elif bad == "max_abs_mz":
        H.max_abs_mz = 101.5
    elif bad == "span":
        H.n_peaks = 4
    assert _abi_call(H, q) == _native.CS_ERR_INVALID
    assert _native.load_cpu().cs_cpu_last_error().decode()


def test_no_gpu_needed():
    """The host engine loads and runs on a machine without a GPU (this one)."""
    from cudams import synthetic
    from cudams.similarity import CPUParallelCosineGreedy

    with warnings.catch_warnings():
        warnings.simplefilter("error")
        p = synthetic.uniform_spectra(8, 3)
        r = CPUParallelCosineGreedy().matrix(as_spectra(p), as_spectra(p))
    np.testing.assert_allclose(np.diag(r["score"]), 1.0, atol=SCORE_ATOL)
