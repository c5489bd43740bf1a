"""Pins the CPU oracle (oracle/cs_oracle.c) before anything is checked against it.

Each check is chosen so a plausible slip in the oracle (window bounds, tie
direction, product order, normalisation, dropped term) fails one of them.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _kats():
    with open(os.path.join(GOLDEN, "cosine_greedy_kats.json")) as fh:
        return json.load(fh)


def compute_expected_score(mz_power, intensity_power, mz1, int1, mz2, int2, matches):
    """Closed form of tests/similarity/test_cosine_greedy.py:9-24 (reference)."""
    mz1, int1, mz2, int2 = map(np.asarray, (mz1, int1, mz2, int2))
    prod = ((mz1[matches[0]] ** mz_power) * (int1[matches[0]] ** intensity_power)
            * (mz2[matches[1]] ** mz_power) * (int2[matches[1]] ** intensity_power))
    den = (np.sqrt((((mz1 ** mz_power) * (int1 ** intensity_power)) ** 2).sum())
           * np.sqrt((((mz2 ** mz_power) * (int2 ** intensity_power)) ** 2).sum()))
    return prod.sum() / den


@pytest.mark.parametrize("case", range(4))
def test_known_answer_pairs(oracle, case):
    k = _kats()["pairs"][case]
    score, acc = oracle.cosine_greedy_assignment(
        k["ref_mz"], k["ref_int"], k["query_mz"], k["query_int"], tolerance=k["tolerance"],
        mz_power=k["mz_power"], intensity_power=k["intensity_power"])
    exp = k["expected_matches"]
    assert sorted(acc) == sorted(zip(exp[0], exp[1]))
    want = compute_expected_score(k["mz_power"], k["intensity_power"], k["ref_mz"], k["ref_int"],
                                  k["query_mz"], k["query_int"], exp)
    assert score == pytest.approx(want, rel=1e-12)


def test_matrix_symmetry(oracle):
    sp = _kats()["matrix"]["spectra"]
    pairs = [(s["mz"], s["int"]) for s in sp]
    res = oracle.cosine_greedy_matrix(pairs, pairs, tolerance=0.1)
    s = res["score"]
    assert s[0, 0] == pytest.approx(s[1, 1], rel=1e-6)
    assert s[0, 1] == pytest.approx(s[1, 0], rel=1e-6)
    assert s[0, 0] == pytest.approx(1.0, rel=1e-12)


def test_pesticides_match_matchms_notebook(oracle):
    from cudams.spectrum import load_from_mgf, normalize_intensities

    with open(os.path.join(GOLDEN, "pesticides_matchms.json")) as fh:
        g = json.load(fh)
    spectra = [normalize_intensities(s) for s in load_from_mgf(os.path.join(GOLDEN, "pesticides.mgf"))]
    assert len(spectra) == 76
    q = spectra[g["query_index"]]
    assert q.get("scans") == g["query_scan"]
    # the notebook prints scores_by_query(sort=True) minus the query itself
    hits = []
    for r in spectra:
        if r is q:
            continue
        s, m, _ = oracle.cosine_greedy_pair(r.peaks.mz, r.peaks.intensities, q.peaks.mz,
                                            q.peaks.intensities, tolerance=g["tolerance"])
        hits.append((s, m, r.get("scans")))
    hits.sort(key=lambda h: -h[0])
    for (s, m, scan), hit in zip(hits, g["hits"]):
        assert scan == hit["scan"], hit
        assert round(s, 4) == pytest.approx(hit["score"], abs=1e-12), hit
        assert m == hit["matches"], hit


def test_committed_pesticide_matrix_matches_oracle(oracle):
    from cudams.spectrum import load_from_mgf, normalize_intensities

    spectra = [normalize_intensities(s) for s in load_from_mgf(os.path.join(GOLDEN, "pesticides.mgf"))]
    pairs = [(s.peaks.mz, s.peaks.intensities) for s in spectra]
    res = oracle.cosine_greedy_matrix(pairs, pairs, tolerance=0.1)
    z = np.load(os.path.join(GOLDEN, "pesticides_oracle.npz"))
    np.testing.assert_array_equal(res["matches"], z["matches"])
    np.testing.assert_array_equal(res["n_candidates"], z["candidates"])
    np.testing.assert_allclose(res["score"], z["score"], rtol=0, atol=0)


def _random_case(rng, n_max=6, grid=0.25, tol=0.3):
    n1, n2 = rng.integers(0, n_max + 1, size=2)
    mz1 = np.sort(100 + grid * rng.integers(0, 12, n1))
    mz2 = np.sort(100 + grid * rng.integers(0, 12, n2))
    i1 = rng.integers(1, 4, n1) / 4.0  # few distinct values: many exact ties
    i2 = rng.integers(1, 4, n2) / 4.0
    return mz1, i1, mz2, i2


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_tiny_inputs(oracle, seed):
    """Direct double loop + explicit (v desc, i desc, j desc) greedy vs the
    stable-sort-then-reverse of the C oracle, on inputs dense in ties and in
    peaks exactly at the tolerance boundary (grid 0.25, tol 0.25/0.5)."""
    rng = np.random.default_rng(seed)
    for _ in range(400):
        mz1, i1, mz2, i2 = _random_case(rng)
        for tol in (0.25, 0.5):
            for mzp, inp in ((0.0, 1.0), (0.5, 2.0)):
                a = oracle.cosine_greedy_pair(mz1, i1, mz2, i2, tolerance=tol, mz_power=mzp,
                                              intensity_power=inp)
                b = oracle.greedy_pair_py(mz1, i1, mz2, i2, tolerance=tol, mz_power=mzp,
                                          intensity_power=inp)
                assert a[1] == b[1] and a[2] == b[2]
                assert a[0] == pytest.approx(b[0], rel=1e-12, abs=1e-300)


def test_plant_flipped_tie_order_is_caught(oracle):
    """Equal values go to the later-collected candidate first.  On this input
    the opposite tie direction accepts a different set, so an oracle with the
    tie order flipped would fail here."""
    mz1, i1 = [101.5, 102.0], [0.5, 0.5]
    mz2, i2 = [102.0, 102.5], [1.0, 0.5]
    s, m, c = oracle.cosine_greedy_pair(mz1, i1, mz2, i2, tolerance=0.6)
    assert (m, c) == (1, 3)
    norm = np.sqrt(0.5) * np.sqrt(1.25)
    assert s == pytest.approx(0.5 / norm, rel=1e-15)

    def flipped(mz1, i1, mz2, i2, tol):  # (v desc, i asc, j asc): the planted mistake
        c = [(i1[a] * i2[b], a, b) for a in range(len(mz1)) for b in range(len(mz2))
             if mz1[a] - tol <= mz2[b] <= mz1[a] + tol]
        c.sort(key=lambda t: (-t[0], t[1], t[2]))
        u1, u2, n = set(), set(), 0
        for _, a, b in c:
            if a not in u1 and b not in u2:
                u1.add(a), u2.add(b)
                n += 1
        return n

    assert flipped(mz1, i1, mz2, i2, 0.6) != m


def test_window_bounds_inclusive(oracle):
    """|mz_q - mz_r| == tol is a match (find_matches uses > / < to exclude)."""
    s, m, c = oracle.cosine_greedy_pair([100.0], [1.0], [100.5], [1.0], tolerance=0.5)
    assert (m, c) == (1, 1)
    s, m, c = oracle.cosine_greedy_pair([100.0], [1.0], [100.5 + 2 ** -40], [1.0], tolerance=0.5)
    assert (m, c) == (0, 0) and s == 0.0


def test_shift_moves_the_window(oracle):
    s, m, c = oracle.cosine_greedy_pair([100.0], [1.0], [90.0], [1.0], tolerance=0.1, shift=10.0)
    assert (m, c) == (1, 1) and s == pytest.approx(1.0)


def test_closed_form_well_separated(oracle):
    """Peaks further apart than 2 tol pair up one-to-one: the score is the
    cosine of the aligned weight vectors."""
    rng = np.random.default_rng(3)
    mz = np.sort(rng.choice(np.arange(100, 1000, 1.0), 40, replace=False))
    a = rng.random(40) + 0.1
    b = rng.random(40) + 0.1
    s, m, _ = oracle.cosine_greedy_pair(mz, a, mz + 0.01, b, tolerance=0.05)
    assert m == 40
    assert s == pytest.approx(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), rel=1e-12)


def test_self_similarity_is_one(oracle):
    from cudams import synthetic

    p = synthetic.lognormal_spectra(20, 11)
    for x in range(p.n_spectra):
        sl = slice(p.offsets[x], p.offsets[x + 1])
        s, m, _ = oracle.cosine_greedy_pair(p.mz[sl], p.intensities[sl], p.mz[sl],
                                            p.intensities[sl], tolerance=0.01)
        assert m == p.offsets[x + 1] - p.offsets[x]
        assert s == pytest.approx(1.0, rel=1e-12)


def test_overflow_rule(oracle):
    mz = [100.0, 100.01, 100.02]
    res = oracle.cosine_greedy_matrix([(mz, [1, 1, 1])], [(mz, [1, 1, 1])], tolerance=0.05,
                                      match_limit=9)
    assert res["n_candidates"][0, 0] == 9 and res["overflow"][0, 0] == 1
    res = oracle.cosine_greedy_matrix([(mz, [1, 1, 1])], [(mz, [1, 1, 1])], tolerance=0.05,
                                      match_limit=10)
    assert res["overflow"][0, 0] == 0


def test_empty_spectra(oracle):
    res = oracle.cosine_greedy_matrix([([], []), ([100.0], [1.0])], [([100.0], [1.0]), ([], [])])
    np.testing.assert_array_equal(res["matches"], [[0, 0], [1, 0]])
    np.testing.assert_array_equal(res["score"], [[0, 0], [1, 0]])


# ----------------------------------------------------------------- fingerprints

def _fkats():
    with open(os.path.join(GOLDEN, "fingerprint_kats.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("measure", ["cosine", "jaccard", "dice"])
def test_fingerprint_known_answers(oracle, measure):
    k = _fkats()
    p = k["pair"]
    s, _ = oracle.fingerprint_matrix(np.array([p["fp1"]]), np.array([p["fp2"]]), measure)
    assert s[0, 0] == pytest.approx(p["expected"][measure], rel=p["rel"])
    e = k["empty"]
    fps = np.array([e["fp1"], e["fp2"]])
    s, _ = oracle.fingerprint_matrix(fps, fps, measure)
    np.testing.assert_allclose(s, e["expected_matrix"], rtol=e["rel"])
    mcase = [c for c in k["missing"]["cases"] if c["measure"] == measure][0]
    fps = np.array([k["missing"]["fp1"], k["missing"]["fp2"]])
    s, _ = oracle.fingerprint_matrix(fps, fps, measure)
    np.testing.assert_allclose(s, [[1, mcase["expected"]], [mcase["expected"], 1]], rtol=1e-7)


def test_fingerprint_counts_brute_force(oracle):
    rng = np.random.default_rng(5)
    a = (rng.random((7, 67)) < 0.3).astype(np.uint8)
    b = (rng.random((5, 67)) < 0.3).astype(np.uint8)
    s, inter = oracle.fingerprint_matrix(a, b, "jaccard")
    for x in range(7):
        for y in range(5):
            i = int(np.sum(a[x] & b[y]))
            u = int(np.sum(a[x] | b[y]))
            assert inter[x, y] == i
            assert s[x, y] == (i / u if u else 0.0)
    s, _ = oracle.fingerprint_matrix(a, b, "dice")
    assert s[2, 3] == pytest.approx(2 * np.sum(a[2] & b[3]) / (a[2].sum() + b[3].sum()))
    s, _ = oracle.fingerprint_matrix(a, b, "cosine")
    assert s[4, 1] == pytest.approx(np.sum(a[4] & b[1]) / np.sqrt(a[4].sum() * b[1].sum()))


def test_near_tie_pairs_constructed(oracle):
    """Two candidates share query peak 0.  Values 0.3 + 1e-12 and 0.3 differ in
    float64 but round to the same float32: the float64 order accepts the first
    collected, the float32 order breaks the tie towards the later one."""
    q = ([100.0], [1.0])
    near = ([100.0, 100.004], [0.3 + 1e-12, 0.3])
    far = ([100.0, 100.004], [0.4, 0.3])
    later_larger = ([100.0, 100.004], [0.3, 0.3 + 1e-12])
    flags = oracle.near_tie_pairs([near, far, later_larger], [q], tolerance=0.01)
    assert flags[:, 0].tolist() == [1, 0, 0]


@pytest.mark.parametrize("seed", range(6))
def test_near_tie_pairs_brute_force(oracle, seed):
    """Against a pure-Python greedy run once with float64 keys and once with
    keys rounded to float32, on tiny inputs whose intensities are drawn from a
    few values a few float32 ulps apart (many near ties)."""
    rng = np.random.default_rng(100 + seed)
    base = np.array([0.3, 0.3 + 1e-12, 0.3 + 3e-12, 0.7, 0.7 - 2e-13])
    spectra = []
    for _ in range(10):
        n = int(rng.integers(1, 6))
        mz = np.sort(rng.choice(np.arange(100.0, 102.0, 0.25), n, replace=False))
        spectra.append((mz, rng.choice(base, n)))
    flags = oracle.near_tie_pairs(spectra, spectra, tolerance=0.3)
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    for r, (m1, i1) in enumerate(spectra):
        for c, (m2, i2) in enumerate(spectra):
            a = oracle.greedy_accepted_py(m1, i1, m2, i2, 0.3)
            b = oracle.greedy_accepted_py(m1, i1, m2, i2, 0.3, key=f32)
            assert flags[r, c] == int(a != b), (r, c)
    assert flags.sum() > 0
