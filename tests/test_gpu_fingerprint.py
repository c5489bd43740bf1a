"""Fingerprint similarity on the GPU (fp4 bit-plane MFMA) against the oracle.  Bar:
intersection counts bit-exact; jaccard/dice bit-exact float32 of the float64
quotient; cosine within 1e-6."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def spectra_with(fps):
    from cudams.spectrum import Spectrum

    out = []
    for fp in fps:
        meta = {} if fp is None else {"fingerprint": np.asarray(fp)}
        out.append(Spectrum(np.array([], float), np.array([], float), meta))
    return out


def _kats():
    with open(os.path.join(GOLDEN, "fingerprint_kats.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("measure", ["cosine", "jaccard", "dice"])
def test_pair_known_answers(gpu, measure):
    """tests/similarity/test_fingerprint_similarity.py:8-32 of the reference."""
    from cudams.similarity import CudaFingerprintSimilarity

    k = _kats()["pair"]
    s1, s2 = spectra_with([k["fp1"], k["fp2"]])
    score = CudaFingerprintSimilarity(similarity_measure=measure).pair(s1, s2)
    assert score == pytest.approx(k["expected"][measure], 1e-6)


@pytest.mark.parametrize("measure", ["cosine", "jaccard", "dice"])
def test_empty_fingerprint(gpu, measure):
    """:35-58 -- an all-zero fingerprint scores 0 (nan_to_num)."""
    from cudams.similarity import CudaFingerprintSimilarity

    k = _kats()["empty"]
    spectra = spectra_with([k["fp1"], k["fp2"]])
    m = CudaFingerprintSimilarity(similarity_measure=measure).matrix(spectra, spectra)
    assert m == pytest.approx(np.array(k["expected_matrix"]), 0.001)


@pytest.mark.parametrize("case", range(3))
def test_missing_fingerprint(gpu, case):
    """:61-112 -- a spectrum without fingerprint gets set_empty_scores."""
    from cudams.similarity import CudaFingerprintSimilarity

    k = _kats()["missing"]
    c = k["cases"][case]
    empty = np.nan if c["empty"] == "nan" else c["empty"]
    spectra = spectra_with([None, k["fp1"], k["fp2"]])
    sim = CudaFingerprintSimilarity(set_empty_scores=empty, similarity_measure=c["measure"])
    m = sim.matrix(spectra, spectra, array_type=c["array_type"])
    if c["array_type"] == "sparse":
        m = m.to_array()
    e = c["expected"]
    want = np.array([[empty, empty, empty], [empty, 1, e], [empty, e, 1]])
    assert np.allclose(m, want, equal_nan=True)


@pytest.mark.parametrize("measure,n_bits,pad", [
    ("jaccard", 1000, 3), ("dice", 1000, 3), ("cosine", 1000, 3),
    # 10 words (one partial 512-bit step), 80 words (odd step count): the
    # bounded-load path; pad 1 makes the row stride a multiple of 4 floats,
    # so the 16-byte stores run with a ragged last group of 4 queries.
    ("jaccard", 300, 1), ("jaccard", 2560, 1), ("dice", 2560, 3),
])
def test_random_vs_oracle(gpu, oracle, measure, n_bits, pad):
    """Ragged shapes (not multiples of the 128 tile, bits not a multiple of 64)."""
    import torch

    from cudams import synthetic
    from cudams.device import fingerprint_similarity

    a = synthetic.fingerprints(300, 3, n_bits=n_bits, zero_rows=3)
    b = synthetic.fingerprints(211, 4, n_bits=n_bits, zero_rows=2)
    wa = torch.from_numpy(synthetic.pack_bits(a).view(np.int32)).cuda()
    wb = torch.from_numpy(synthetic.pack_bits(b).view(np.int32)).cuda()
    out = torch.full((300, 211 + pad), -7.0, dtype=torch.float32, device="cuda")
    inter = torch.full((300, 211 + pad), -7, dtype=torch.int32, device="cuda")
    fingerprint_similarity(measure, wa, wb, out, intersections=inter)
    torch.cuda.synchronize()
    want, wi = oracle.fingerprint_matrix(a, b, measure)
    np.testing.assert_array_equal(inter.cpu().numpy()[:, :211], wi)
    # columns past the last query are left alone
    assert (out.cpu().numpy()[:, 211:] == -7.0).all()
    assert (inter.cpu().numpy()[:, 211:] == -7).all()
    got = out.cpu().numpy()[:, :211]
    if measure == "cosine":
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    else:
        np.testing.assert_array_equal(got, want.astype(np.float32))


def test_full_size_sampled(gpu, oracle):
    """4096 x 4096 x 2048 bits (density U[0.02, 0.3]): intersections of 128
    random reference rows against every query, exact."""
    import torch

    from cudams import synthetic
    from cudams.device import fingerprint_similarity

    a = synthetic.fingerprints(4096, 3)
    b = synthetic.fingerprints(4096, 4)
    wa = torch.from_numpy(synthetic.pack_bits(a).view(np.int32)).cuda()
    wb = torch.from_numpy(synthetic.pack_bits(b).view(np.int32)).cuda()
    out = torch.empty((4096, 4096), dtype=torch.float32, device="cuda")
    inter = torch.empty((4096, 4096), dtype=torch.int32, device="cuda")
    fingerprint_similarity("jaccard", wa, wb, out, intersections=inter)
    torch.cuda.synchronize()
    rows = np.random.default_rng(1).choice(4096, 128, replace=False)
    want, wi = oracle.fingerprint_matrix(a[rows], b, "jaccard")
    np.testing.assert_array_equal(inter.cpu().numpy()[rows], wi)
    np.testing.assert_array_equal(out.cpu().numpy()[rows], want.astype(np.float32))
    # symmetry of the intersection count under swapping the operands
    inter2 = torch.empty((4096, 4096), dtype=torch.int32, device="cuda")
    fingerprint_similarity("jaccard", wb, wa, out, intersections=inter2)
    torch.cuda.synchronize()
    assert torch.equal(inter, inter2.t())


@pytest.mark.parametrize("measure", ["jaccard", "dice", "cosine"])
def test_full_size_every_pair(gpu, oracle, measure):
    """The whole 4096 x 4096 x 2048-bit workload (seeds 3 and 4) for each
    measure against the oracle over every pair: intersections exact,
    jaccard and dice the float32 of the float64 quotient, cosine within
    1e-6."""
    import os as _os

    import torch

    from cudams import synthetic
    from cudams.device import fingerprint_similarity

    a = synthetic.fingerprints(4096, 3)
    b = synthetic.fingerprints(4096, 4)
    wa = torch.from_numpy(synthetic.pack_bits(a).view(np.int32)).cuda()
    wb = torch.from_numpy(synthetic.pack_bits(b).view(np.int32)).cuda()
    out = torch.empty((4096, 4096), dtype=torch.float32, device="cuda")
    inter = torch.empty((4096, 4096), dtype=torch.int32, device="cuda")
    print(f"fingerprints 4096x4096x2048 {measure}", flush=True)
    fingerprint_similarity(measure, wa, wb, out, intersections=inter)
    torch.cuda.synchronize()
    want, wi = oracle.fingerprint_matrix(a, b, measure, nthreads=len(_os.sched_getaffinity(0)))
    np.testing.assert_array_equal(inter.cpu().numpy(), wi)
    got = out.cpu().numpy()
    if measure == "cosine":
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    else:
        np.testing.assert_array_equal(got, want.astype(np.float32))


@pytest.mark.parametrize("measure", ["jaccard", "dice", "cosine"])
def test_every_count_pair_up_to_4096_bits(gpu, measure):
    """Fingerprints of 4096 bits whose set bits are prefixes: reference a and
    query b (a, b = 0..4096) intersect in min(a, b) bits, so the 4097 x 4097
    matrix holds every (intersection, denominator) pair the measures can meet
    at this length: jaccard min/max, dice 2 min/(a + b), every n <= d.  The
    division the kernel uses up to 4096 bits must give the float32 rounding of
    the float64 quotient on all of them (bit for bit); cosine is within
    1e-6."""
    import torch

    from cudams import synthetic
    from cudams.device import fingerprint_similarity

    n, bits = 4097, 4096
    fp = (np.arange(bits)[None, :] < np.arange(n)[:, None]).astype(np.uint8)
    w = torch.from_numpy(synthetic.pack_bits(fp).view(np.int32)).cuda()
    out = torch.empty((n, n), dtype=torch.float32, device="cuda")
    inter = torch.empty((n, n), dtype=torch.int32, device="cuda")
    print(f"prefix fingerprints {n}x{n}x{bits} {measure}", flush=True)
    fingerprint_similarity(measure, w, w, out, intersections=inter)
    torch.cuda.synchronize()
    a = np.arange(n, dtype=np.float64)
    lo = np.minimum.outer(a, a)
    np.testing.assert_array_equal(inter.cpu().numpy(), lo.astype(np.int32))
    with np.errstate(invalid="ignore", divide="ignore"):
        if measure == "jaccard":
            want = lo / np.maximum.outer(a, a)
        elif measure == "dice":
            want = 2.0 * (lo / np.add.outer(a, a))
        else:
            want = lo / np.sqrt(np.multiply.outer(a, a))
    want = np.nan_to_num(want)
    got = out.cpu().numpy()
    if measure == "cosine":
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    else:
        bad = got != want.astype(np.float32)
        assert int(bad.sum()) == 0, f"{int(bad.sum())} pairs differ, e.g. {np.argwhere(bad)[:5]}"
