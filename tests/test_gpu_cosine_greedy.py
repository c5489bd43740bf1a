"""CosineGreedy on the GPU (libcsamd.so through the drop-in class and the
device API) against the oracle.  Bar: matches / candidates-derived overflow
bit-exact, score within 1e-5 absolute of the float64 oracle."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SCORE_ATOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spectra_list(p):
    from cudams.spectrum import Spectrum

    return [Spectrum(p.mz[p.offsets[i]:p.offsets[i + 1]],
                     p.intensities[p.offsets[i]:p.offsets[i + 1]]) for i in range(p.n_spectra)]


def assert_parity(res, want, match_limit=2048):
    np.testing.assert_array_equal(res["matches"], want["matches"])
    np.testing.assert_array_equal(res["overflow"], (want["n_candidates"] >= match_limit).astype(np.uint8))
    np.testing.assert_allclose(res["score"], want["score"], rtol=0, atol=SCORE_ATOL)


def device_run(refs, queries, tol=0.1, shift=0.0, mzp=0.0, inp=1.0, match_limit=2048, K=0, T=0,
               ld_pad=0):
    import torch

    from cudams.device import cosine_greedy
    from cudams.spectra import DeviceSpectra

    R = DeviceSpectra(refs, mzp, inp)
    Q = DeviceSpectra(queries, mzp, inp)
    ld = Q.n_spectra + ld_pad
    score = torch.full((R.n_spectra, ld), -7.0, dtype=torch.float32, device="cuda")
    matches = torch.full((R.n_spectra, ld), -7, dtype=torch.int32, device="cuda")
    overflow = torch.full((R.n_spectra, ld), 7, dtype=torch.uint8, device="cuda")
    cosine_greedy(R, Q, tol, shift, match_limit, score, matches, overflow, light_candidates=K,
                  round_capacity=T)
    torch.cuda.synchronize()
    out = {"score": score.cpu().numpy(), "matches": matches.cpu().numpy(),
           "overflow": overflow.cpu().numpy()}
    if ld_pad:
        assert np.all(out["score"][:, Q.n_spectra:] == -7.0)
        assert np.all(out["matches"][:, Q.n_spectra:] == -7)
        out = {k: v[:, :Q.n_spectra] for k, v in out.items()}
    return out


def oracle_csr(oracle, refs, queries, **kw):
    return oracle.cosine_greedy_matrix((refs.mz, refs.intensities, refs.offsets),
                                       (queries.mz, queries.intensities, queries.offsets), **kw)


@pytest.mark.parametrize("case", range(4))
def test_known_answer_pairs(gpu, case):
    """tests/similarity/test_cosine_greedy.py:27-120 of the reference, through pair()."""
    from cudams.similarity import CudaCosineGreedy
    from cudams.spectrum import Spectrum

    with open(os.path.join(GOLDEN, "cosine_greedy_kats.json")) as fh:
        k = json.load(fh)["pairs"][case]
    say(f"KAT case {case}")
    s1 = Spectrum(np.array(k["ref_mz"], float), np.array(k["ref_int"], float))
    s2 = Spectrum(np.array(k["query_mz"], float), np.array(k["query_int"], float))
    cg = CudaCosineGreedy(batch_size=8, tolerance=k["tolerance"], mz_power=k["mz_power"],
                          intensity_power=k["intensity_power"])
    score = cg.pair(s1, s2)
    exp = k["expected_matches"]
    mz1, i1 = np.array(k["ref_mz"], float), np.array(k["ref_int"], float)
    mz2, i2 = np.array(k["query_mz"], float), np.array(k["query_int"], float)
    p, q = k["mz_power"], k["intensity_power"]
    num = ((mz1[exp[0]] ** p) * (i1[exp[0]] ** q) * (mz2[exp[1]] ** p) * (i2[exp[1]] ** q)).sum()
    den = np.sqrt((((mz1 ** p) * (i1 ** q)) ** 2).sum()) * np.sqrt((((mz2 ** p) * (i2 ** q)) ** 2).sum())
    assert score["score"] == pytest.approx(num / den, 0.0001)
    assert score["matches"] == len(exp[0])


@pytest.mark.parametrize("symmetric", [True, False])
def test_matrix_symmetry(gpu, symmetric):
    from cudams.similarity import CudaCosineGreedy
    from cudams.spectrum import Spectrum

    s1 = Spectrum(np.array([100, 200, 300], float), np.array([0.1, 0.2, 1.0], float))
    s2 = Spectrum(np.array([110, 190, 290], float), np.array([0.5, 0.2, 1.0], float))
    spectra = [s1, s2]
    scores = CudaCosineGreedy(batch_size=8).matrix(spectra, spectra, is_symmetric=symmetric)
    assert scores.dtype.names == ("score", "matches", "overflow")
    assert scores[0][0]["score"] == pytest.approx(scores[1][1]["score"], 0.000001)
    assert scores[0][1]["score"] == pytest.approx(scores[1][0]["score"], 0.000001)


def test_uniform_256_vs_oracle(gpu, oracle):
    """The 256 x 256 uniform workload (refs seed 1, queries seed 2), tol 0.1."""
    from cudams import synthetic

    refs = synthetic.uniform_spectra(256, 1)
    ques = synthetic.uniform_spectra(256, 2)
    say("uniform 256x256 tol 0.1")
    res = device_run(refs, ques, tol=0.1)
    assert_parity(res, oracle_csr(oracle, refs, ques, tolerance=0.1))


@pytest.mark.parametrize("tol", [0.25, 0.5, 0.1, 0.3, 0.0])
def test_window_edges_on_a_grid(gpu, oracle, tol):
    """m/z on a 0.25 grid (float32 numbers, the float32 kernel variant): with
    tolerance 0.25 or 0.5 many peaks sit exactly on a window edge, where the
    inclusive float64 test decides; 0.1 and 0.3 are not float32 numbers;
    tolerance 0 matches equal m/z only.
    Integer-quarter intensities make equal products, so the tie order counts
    too.  Two query tiles, the second partial."""
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
    say(f"grid edges tol {tol}")
    want = oracle_csr(oracle, refs, ques, tolerance=tol)
    assert want["n_candidates"].max() > (1 if tol == 0.0 else 8)  # wave path above 8
    for K in (1, 8):
        assert_parity(device_run(refs, ques, tol=tol, K=K), want)


def test_uniform_slice_fixture(gpu):
    """Same workload, first 32 x 32, against the committed golden fixture."""
    from cudams import synthetic

    z = np.load(os.path.join(GOLDEN, "uniform_slice.npz"))
    refs = synthetic.uniform_spectra(256, 1)
    ques = synthetic.uniform_spectra(256, 2)
    res = device_run(refs, ques, tol=0.1)
    assert np.array_equal(refs.mz[: len(z["ref_mz"])], z["ref_mz"])
    np.testing.assert_array_equal(res["matches"][:32, :32], z["matches"])
    np.testing.assert_allclose(res["score"][:32, :32], z["score"], atol=SCORE_ATOL, rtol=0)


def test_pesticides_matrix_and_golden(gpu):
    """pesticides.mgf all-vs-all through CudaCosineGreedy.matrix, against the
    committed oracle matrix and the matchms values of the reference notebook."""
    from cudams.similarity import CudaCosineGreedy
    from cudams.spectrum import load_from_mgf, normalize_intensities

    spectra = [normalize_intensities(s) for s in load_from_mgf(os.path.join(GOLDEN, "pesticides.mgf"))]
    say("pesticides 76x76")
    res = CudaCosineGreedy(batch_size=512).matrix(spectra, spectra)
    z = np.load(os.path.join(GOLDEN, "pesticides_oracle.npz"))
    np.testing.assert_array_equal(res["matches"], z["matches"])
    np.testing.assert_allclose(res["score"], z["score"], atol=SCORE_ATOL, rtol=0)
    assert not res["overflow"].any()
    with open(os.path.join(GOLDEN, "pesticides_matchms.json")) as fh:
        g = json.load(fh)
    q = g["query_index"]
    col = [(res["score"][r, q], res["matches"][r, q], spectra[r].get("scans"))
           for r in range(len(spectra)) if r != q]
    col.sort(key=lambda t: -t[0])
    for (s, m, scan), hit in zip(col, g["hits"]):
        assert scan == hit["scan"] and m == hit["matches"]
        assert abs(s - hit["score"]) < 5e-5


@pytest.mark.parametrize("K,T", [(1, 4), (4, 8), (8, 512), (16, 1024)])
def test_heavy_path_forced(gpu, oracle, K, T):
    """Small light capacity and round capacity: most pairs go through the
    wave-cooperative path, many over several rounds."""
    from cudams import synthetic

    refs = synthetic.lognormal_spectra(40, 21, cluster_fraction=0.5)
    ques = synthetic.lognormal_spectra(70, 22, cluster_fraction=0.5)
    for tol in (0.01, 0.5):
        say(f"heavy K={K} T={T} tol={tol}")
        res = device_run(refs, ques, tol=tol, K=K, T=T, match_limit=64)
        assert_parity(res, oracle_csr(oracle, refs, ques, tolerance=tol, match_limit=64), 64)


def test_self_pairs_with_clusters(gpu, oracle):
    """All-vs-all with dense clusters: self pairs carry thousands of candidates."""
    from cudams import synthetic

    p = synthetic.lognormal_spectra(96, 5)
    say("clusters 96x96 tol 0.01")
    res = device_run(p, p, tol=0.01)
    want = oracle_csr(oracle, p, p, tolerance=0.01)
    assert want["n_candidates"].max() > 1000
    assert_parity(res, want)


@pytest.mark.parametrize("mzp,inp,tol", [(0.5, 2.0, 1.0), (1.0, 0.5, 0.1), (0.0, 2.0, 0.02)])
def test_powers(gpu, oracle, mzp, inp, tol):
    from cudams import synthetic

    refs = synthetic.uniform_spectra(64, 31)
    ques = synthetic.uniform_spectra(80, 32)
    say(f"powers {mzp} {inp} tol {tol}")
    res = device_run(refs, ques, tol=tol, mzp=mzp, inp=inp)
    assert_parity(res, oracle_csr(oracle, refs, ques, tolerance=tol, mz_power=mzp,
                                  intensity_power=inp))


def test_float64_mz_and_shift(gpu, oracle):
    """m/z that float32 cannot hold, many within ulps of the window edge, plus
    a nonzero shift and a padded leading dimension."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(9)
    mzs_r, its_r, mzs_q, its_q = [], [], [], []
    for _ in range(50):
        base = np.sort(rng.uniform(100, 900, rng.integers(1, 40)))
        mzs_r.append(base)
        its_r.append(rng.random(len(base)) + 0.01)
        tol = 0.01
        eps = rng.integers(-3, 4, len(base)) * np.spacing(base)
        q = np.sort(base + rng.choice([-tol, tol], len(base)) + eps - 3.5)
        mzs_q.append(q)
        its_q.append(rng.random(len(q)) + 0.01)
    refs = PackedSpectra.from_arrays(mzs_r, its_r)
    ques = PackedSpectra.from_arrays(mzs_q, its_q)
    say("float64 m/z + shift 3.5")
    res = device_run(refs, ques, tol=0.01, shift=3.5, ld_pad=5)
    assert_parity(res, oracle_csr(oracle, refs, ques, tolerance=0.01, shift=3.5))


def test_empty_and_ragged(gpu, oracle):
    from cudams.spectra import PackedSpectra

    mzs = [[], [100.0], [100.0, 100.05, 300.0], [], [99.95]]
    its = [[], [1.0], [0.5, 1.0, 0.2], [], [0.7]]
    p = PackedSpectra.from_arrays(mzs, its)
    say("empty and ragged")
    res = device_run(p, p, tol=0.1)
    assert_parity(res, oracle_csr(oracle, p, p, tolerance=0.1))


@pytest.mark.parametrize("K", [1, 8])
def test_zero_intensities(gpu, oracle, K):
    """Peaks of intensity 0 (candidate value 0: ties broken by (i, j) alone)
    and one all-zero spectrum per side, whose zero norm makes every pair
    with a candidate 0/0 = NaN, as in matchms; dense enough at tol 0.5 for
    the wave and heavy paths."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(41)

    def spectra(n, seed_shift):
        mzs, its = [], []
        for x in range(n):
            k = int(rng.integers(3, 120))
            mz = np.sort(rng.uniform(100, 200, k).astype(np.float32)).astype(np.float64)
            it = rng.random(k) * (rng.random(k) > 0.3)
            if x == seed_shift:
                it = np.zeros(k)
            mzs.append(mz)
            its.append(it)
        return PackedSpectra.from_arrays(mzs, its)

    refs, ques = spectra(21, 4), spectra(70, 9)
    say(f"zero intensities K {K}")
    want = oracle_csr(oracle, refs, ques, tolerance=0.5)
    assert np.isnan(want["score"]).any() and (want["n_candidates"] > 64).any()
    assert_parity(device_run(refs, ques, tol=0.5, K=K), want)


def test_overflow_flag_rule(gpu, oracle):
    from cudams import synthetic

    p = synthetic.lognormal_spectra(48, 8)
    say("overflow rule, match_limit 16")
    res = device_run(p, p, tol=0.01, match_limit=16)
    want = oracle_csr(oracle, p, p, tolerance=0.01, match_limit=16)
    assert want["overflow"].any()
    assert_parity(res, want, 16)


def test_sparse_output(gpu, oracle):
    from cudams import synthetic
    from cudams.similarity import CudaCosineGreedy

    p = synthetic.uniform_spectra(120, 41)
    spectra = spectra_list(p)
    cg = CudaCosineGreedy(tolerance=0.1, batch_size=50, sparse_threshold=0.05)
    say("sparse output")
    sp = cg.matrix(spectra, spectra, array_type="sparse")
    want = oracle_csr(oracle, p, p, tolerance=0.1)
    keep = want["score"] >= 0.05
    dense = sp.to_array("sparse_score")
    assert set(zip(*np.nonzero(keep))) <= set(zip(sp.row.tolist(), sp.col.tolist())) | set(
        zip(*np.nonzero(np.abs(want["score"] - 0.05) < SCORE_ATOL)))
    np.testing.assert_allclose(dense[keep], want["score"][keep], atol=SCORE_ATOL, rtol=0)
    np.testing.assert_array_equal(sp.to_array("sparse_matches")[keep], want["matches"][keep])


def test_large_sampled_pairs(gpu, oracle):
    """2048 x 2048 of the clustered lognormal workload: full-matrix invariants
    plus 4000 random pairs against the oracle."""
    from cudams import synthetic

    p = synthetic.lognormal_spectra(2048, 5)
    say("2048x2048 clustered tol 0.01")
    res = device_run(p, p, tol=0.01)
    n = np.diff(p.offsets)
    assert np.all(np.diag(res["matches"]) == n)
    np.testing.assert_allclose(np.diag(res["score"]), 1.0, atol=SCORE_ATOL)
    assert np.all(res["matches"] >= 0)
    assert np.all(res["matches"] <= np.minimum.outer(n, n))
    assert np.all((res["score"] >= 0) & (res["score"] <= 1 + 1e-6))
    rng = np.random.default_rng(0)
    rr = rng.integers(0, 2048, 4000)
    qq = rng.integers(0, 2048, 4000)
    for r, q in zip(rr, qq):
        a = slice(p.offsets[r], p.offsets[r + 1])
        b = slice(p.offsets[q], p.offsets[q + 1])
        s, m, c = oracle.cosine_greedy_pair(p.mz[a], p.intensities[a], p.mz[b], p.intensities[b],
                                            tolerance=0.01)
        assert res["matches"][r, q] == m
        assert abs(res["score"][r, q] - s) <= SCORE_ATOL
        assert res["overflow"][r, q] == (c >= 2048)


def test_c3_full_matrix_vs_oracle(gpu, oracle):
    """The whole 8192 x 8192 clustered lognormal workload (seed 5, tolerance
    0.01, match_limit 2048) against the oracle: all 67,108,864 pairs, matches
    and overflow bit-exact, score within 1e-5.  Near-tie exposure: the pairs
    whose accepted set changes when candidate values are rounded to float32
    (what the reference GPU kernel stores) must come out as the float64 oracle
    has them."""
    import os as _os

    from cudams import synthetic

    p = synthetic.lognormal_spectra(8192, 5)
    say("8192x8192 clustered tol 0.01 on the GPU")
    res = device_run(p, p, tol=0.01)
    cores = len(_os.sched_getaffinity(0))
    say(f"oracle on {cores} cores")
    csr = (p.mz, p.intensities, p.offsets)
    want = oracle.cosine_greedy_matrix(csr, csr, tolerance=0.01, nthreads=cores)
    say("comparing")
    assert int((res["matches"] != want["matches"]).sum()) == 0
    np.testing.assert_array_equal(res["overflow"], (want["n_candidates"] >= 2048).astype(np.uint8))
    err = np.abs(res["score"].astype(np.float64) - want["score"])
    assert float(err.max()) <= SCORE_ATOL
    # within two float32 ulps of the float64 score: the float32 of the same sum
    ulp = np.spacing(np.abs(want["score"]).astype(np.float32)).astype(np.float64)
    assert int((err > 2 * ulp).sum()) == 0
    ties = oracle.near_tie_pairs(csr, csr, tolerance=0.01, nthreads=cores)
    say(f"near-tie pairs {int(ties.sum())}")
    assert int(want["overflow"].sum()) > 0 and int((want["n_candidates"] > 64).sum()) > 1000


def test_tiles_larger_than_lds(gpu, oracle):
    """Query tiles whose sorted peak list exceeds the LDS capacity take the
    HBM-resident index path; tiles of both kinds in one call."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(12)
    mzs, its = [], []
    for x in range(96):
        k = int(rng.integers(250, 400)) if x < 40 else int(rng.integers(5, 60))
        mzs.append(np.sort(rng.uniform(50, 1000, k).astype(np.float32)).astype(np.float64))
        its.append(rng.random(k) + 0.01)
    ques = PackedSpectra.from_arrays(mzs, its)
    refs = PackedSpectra.from_arrays(mzs[::3], its[::3])
    say("tiles beyond LDS, tol 0.05")
    res = device_run(refs, ques, tol=0.05)
    assert_parity(res, oracle_csr(oracle, refs, ques, tolerance=0.05))


@pytest.mark.parametrize("tol", [0.05, 0.5])
def test_spectra_above_staging_size(gpu, oracle, tol):
    """Spectra of 1100-1400 peaks (above the 1024 the cooperative kernel
    stages in LDS): its enumeration reads the spectra from HBM and its greedy
    runs without the staging-area tables.  Tolerance 0.05 gives hundreds of
    candidates per pair (heavy kernel), 0.5 thousands (giant kernel too)."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(23)

    def spectra(n):
        mzs, its = [], []
        for _ in range(n):
            k = int(rng.integers(1100, 1400))
            mzs.append(np.sort(rng.uniform(100, 300, k).astype(np.float32)).astype(np.float64))
            its.append(rng.random(k) + 0.01)
        return PackedSpectra.from_arrays(mzs, its)

    refs, ques = spectra(5), spectra(7)
    assert max(np.diff(refs.offsets).max(), np.diff(ques.offsets).max()) > 1024
    say(f"spectra above staging size, tol {tol}")
    want = oracle_csr(oracle, refs, ques, tolerance=tol)
    assert want["n_candidates"].min() > 64  # every pair leaves the lane kernel
    if tol == 0.5:
        assert want["n_candidates"].max() > 4096  # some go on to the giant kernel
    assert_parity(device_run(refs, ques, tol=tol), want)


def test_stage_reuse_with_other_references(gpu, oracle):
    """PREPARE once for a query set, then LIGHT|HEAVY with two different
    reference sets on the same workspace."""
    import torch

    from cudams import _native, synthetic
    from cudams.device import Workspace, cosine_greedy
    from cudams.spectra import DeviceSpectra

    lib = synthetic.lognormal_spectra(300, 61)
    Q = DeviceSpectra(lib)
    ws = Workspace("cuda")
    first = True
    for seed, n in ((62, 17), (63, 130)):
        refs = synthetic.lognormal_spectra(n, seed)
        R = DeviceSpectra(refs)
        s = torch.empty((n, 300), dtype=torch.float32, device="cuda")
        m = torch.empty((n, 300), dtype=torch.int32, device="cuda")
        o = torch.empty((n, 300), dtype=torch.uint8, device="cuda")
        if first:
            cosine_greedy(R, Q, 0.01, 0.0, 2048, s, m, o, ws, stages=_native.CS_STAGE_PREPARE)
            first = False
        cosine_greedy(R, Q, 0.01, 0.0, 2048, s, m, o, ws,
                      stages=_native.CS_STAGE_LIGHT | _native.CS_STAGE_HEAVY)
        torch.cuda.synchronize()
        res = {"score": s.cpu().numpy(), "matches": m.cpu().numpy(), "overflow": o.cpu().numpy()}
        assert_parity(res, oracle_csr(oracle, refs, lib, tolerance=0.01))


def test_heavy_sort_widths(gpu, oracle):
    """Pairs whose candidate counts fall in every padded sort size of the
    heavy kernel (512, 1024, 2048, 4096 entries: register passes of 2, 4, 8
    and 16 entries per thread) and below: dense peak packs inside a wide
    tolerance give n_r x n_q candidates, with tied intensities."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(91)
    mzs, its = [], []
    for k in (12, 18, 22, 30, 34, 42, 48, 55, 60):
        pack = 300.0 + rng.uniform(0, 0.8, k).astype(np.float32)
        rest = rng.uniform(400.0, 900.0, 20).astype(np.float32)
        mzs.append(np.sort(np.concatenate([pack, rest])).astype(np.float64))
        its.append(np.round(rng.random(k + 20) * 8 + 1) / 8)  # ties
    p = PackedSpectra.from_arrays(mzs, its)
    say("heavy sort widths")
    res = device_run(p, p, tol=1.0, match_limit=5000)
    want = oracle_csr(oracle, p, p, tolerance=1.0, match_limit=5000)
    c = want["n_candidates"]
    for lo, hi in ((65, 256), (257, 512), (513, 1024), (1025, 2048), (2049, 4096)):
        assert ((c >= lo) & (c <= hi)).any(), (lo, hi)
    assert c.max() <= 4096
    assert_parity(res, want, 5000)


def heavy_near_tie_gadget(n_pack_r, n_pack_q, seed):
    """References and queries whose every pair is heavy (a dense pack of
    n_pack_r x n_pack_q candidates near m/z 500, tolerance 1) and holds a
    three-candidate path (a0, b0), (a0, b1), (a1, b1) near m/z 100 with values
    1 + 2^-40, (1 + 2^-40)(1 + 2^-41), 1 + 2^-41.  The sequential greedy
    takes (a0, b1) first and matches one peak of the path; taken in (i, j)
    order -- the order of equal 44-bit value prefixes in the heavy kernel's
    64-bit keys -- it would match two."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(seed)
    refs_mz, refs_in, q_mz, q_in = [], [], [], []
    for k in n_pack_r:
        pack = np.sort(500.0 + rng.uniform(0, 0.8, k).astype(np.float32)).astype(np.float64)
        refs_mz.append(np.concatenate([[100.0, 101.5], pack]))
        refs_in.append(np.concatenate([[1.0 + 2.0 ** -40, 1.0],
                                       np.round(rng.uniform(0.1, 0.6, k), 6)]))
    for k in n_pack_q:
        pack = np.sort(500.0 + rng.uniform(0, 0.8, k).astype(np.float32)).astype(np.float64)
        q_mz.append(np.concatenate([[99.5, 100.8], pack]))
        q_in.append(np.concatenate([[1.0, 1.0 + 2.0 ** -41], np.round(rng.uniform(0.1, 0.6, k), 6)]))
    return PackedSpectra.from_arrays(refs_mz, refs_in), PackedSpectra.from_arrays(q_mz, q_in)


def test_heavy_values_below_key_resolution(gpu, oracle):
    """Heavy pairs whose greedy depends on candidate values that differ only
    below the heavy kernel's 44-bit key prefix (heavy_near_tie_gadget): the
    kernel's exact check of equal-prefix neighbours must settle them in value
    order.  Pack sizes cover the one-wave sort (<= 256 candidates) and the
    workgroup sort."""
    refs, queries = heavy_near_tie_gadget((9, 12, 20, 33), (8, 11, 16, 30), 57)
    say("heavy values below key resolution")
    res = device_run(refs, queries, tol=1.0, match_limit=5000)
    want = oracle_csr(oracle, refs, queries, tolerance=1.0, match_limit=5000)
    c = want["n_candidates"]
    assert c.min() > 64 and (c <= 256).any() and (c > 512).any()
    # the path matches one peak in every pair: the pack's matches plus one
    packs_r, packs_q = np.diff(refs.offsets) - 2, np.diff(queries.offsets) - 2
    np.testing.assert_array_equal(want["matches"], np.minimum.outer(packs_r, packs_q) + 1)
    assert_parity(res, want, 5000)


@pytest.mark.parametrize("T", [4, 64, 1024])
def test_giant_pairs(gpu, oracle, T):
    """Pairs with more candidates than the single-pass buffer (4096): dense
    peak packs inside a wide tolerance give 5k-40k candidates per pair; the
    round-based path settles them T at a time."""
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(77)
    mzs, its = [], []
    for x in range(9):
        k = int(rng.integers(60, 200))
        base = 300.0 + 40.0 * (x % 3)
        mzs.append(np.sort(base + rng.uniform(0, 0.8, k).astype(np.float32)).astype(np.float64))
        its.append(np.round(rng.random(k) * 8 + 1) / 8)  # ties
    p = PackedSpectra.from_arrays(mzs, its)
    say(f"giant pairs T={T}")
    res = device_run(p, p, tol=1.0, T=T, match_limit=5000)
    want = oracle_csr(oracle, p, p, tolerance=1.0, match_limit=5000)
    assert want["n_candidates"].max() > 4096
    assert_parity(res, want, 5000)


@pytest.mark.parametrize("n,k", [(1040, 70), (730, 270)])
def test_heavy_queue_overflow(gpu, n, k):
    """More heavy pairs than their region of the 2^20-entry queue holds: n
    spectra whose k peaks all sit at the same m/z (1 Da apart, tol 0.1), so
    every pair has exactly the k diagonal candidates (above the 64 the lane
    kernel settles).  k = 70: pairs settled one per wave (at most 256
    candidates and peaks; their region holds 2^19), k = 270: pairs settled by
    a workgroup each (their size class's region holds 2^17).  The overflowing
    queue is abandoned for a scan of the matches plane; the closed form is
    matches = k and score = cosine of the intensity vectors."""
    from cudams.spectra import PackedSpectra

    assert n * n > (1 << 19 if k <= 256 else 1 << 17)
    rng = np.random.default_rng(21)
    mz = 100.0 + np.arange(k, dtype=np.float64)
    its = rng.random((n, k)).astype(np.float32).astype(np.float64) + 0.05
    p = PackedSpectra.from_arrays([mz] * n, list(its))
    say(f"heavy queue overflow {n}x{n}")
    res = device_run(p, p, tol=0.1)
    assert np.all(res["matches"] == k)
    w = p.intensities.reshape(n, k)
    norm = np.sqrt((w * w).sum(axis=1))
    want = (w @ w.T) / np.outer(norm, norm)
    np.testing.assert_allclose(res["score"], want, rtol=0, atol=SCORE_ATOL)
    assert np.all(res["overflow"] == 0)
