"""Host-side logic that needs no GPU: CSR packing and validation, power
factors, MGF reading, synthetic generators, bit packing, the sparse container
and the C-ABI library's exported symbols."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO


def test_library_exports_every_declared_symbol():
    """Every function include/cosine_similarity_amd.h declares is exported by
    libcsamd.so (loaded, not called: no GPU here)."""
    import re

    from cudams import _native

    L = _native.load()
    C = _native.load_cpu()
    with open(os.path.join(REPO, "include", "cosine_similarity_amd.h")) as fh:
        hdr = fh.read()
    declared = set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_native.EXPORTED_SYMBOLS) | set(_native.CPU_EXPORTED_SYMBOLS)
    for name in _native.EXPORTED_SYMBOLS:
        assert hasattr(L, name), name
    for name in _native.CPU_EXPORTED_SYMBOLS:
        assert hasattr(C, name), name
    assert L.cs_version() >= 10000
    assert L.cs_last_error() == b""


def test_library_is_gfx950_code_object():
    from cudams import _native

    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-n", _native.LIB_PATH],
                         capture_output=True, text=True)
    blob = open(_native.LIB_PATH, "rb").read()
    assert b"gfx950" in blob or "gfx950" in out.stdout


def test_struct_layouts_match_header():
    from cudams import _native

    assert ctypes.sizeof(_native.CsSpectra) == 8 * 8 + 4 + 4
    assert ctypes.sizeof(_native.CsCosineParams) == 8 + 8 + 8 + 4 + 4 + 4 + 4


def test_invalid_arguments_rejected_without_gpu():
    """Argument checks run before any HIP call, so they work here."""
    from cudams import _native

    L = _native.load()
    n = ctypes.c_size_t()
    bad = _native.CsSpectra(1, 1, None, None, None, None, None, None, 70000, 1.0)
    ok = _native.CsSpectra(0, 0, None, None, None, None, None, None, 0, 0.0)
    st = L.cs_cosine_greedy_workspace_size(ctypes.byref(bad), ctypes.byref(ok), ctypes.byref(n))
    assert st == _native.CS_ERR_UNSUPPORTED
    assert b"max_peaks" in L.cs_last_error()
    with pytest.raises(ValueError):
        _native.check(L.cs_cosine_greedy_workspace_size(None, ctypes.byref(ok), ctypes.byref(n)))
    st = L.cs_fingerprint_similarity(7, None, 1, None, 1, 1, None, None, 0.0, None, None, 1,
                                     None, 0, None)
    assert st == _native.CS_ERR_INVALID


def test_pack_truncates_and_validates():
    from cudams.spectra import PackedSpectra
    from cudams.spectrum import Spectrum

    s = [Spectrum(np.array([1.0, 2.0, 3.0]), np.array([1.0, 0.5, 0.2])),
         Spectrum(np.array([]), np.array([])), None, (np.array([5.0]), np.array([1.0]))]
    p = PackedSpectra.from_spectra(s, n_max_peaks=2)
    assert p.offsets.tolist() == [0, 2, 2, 2, 3]
    assert p.mz.tolist() == [1.0, 2.0, 5.0]
    with pytest.raises(ValueError):
        PackedSpectra.from_arrays([[2.0, 1.0]], [[1.0, 1.0]])
    with pytest.raises(ValueError):
        PackedSpectra.from_arrays([[1.0, np.nan]], [[1.0, 1.0]])
    with pytest.raises(ValueError):
        PackedSpectra.from_arrays([np.arange(70000.0)], [np.ones(70000)])
    # descending across a spectrum boundary is fine
    PackedSpectra.from_arrays([[5.0, 6.0], [1.0, 2.0]], [[1, 1], [1, 1]])


def test_power_factors_match_matchms_expression():
    from cudams.spectra import PackedSpectra, power_factors

    rng = np.random.default_rng(0)
    mzs = [np.sort(rng.uniform(50, 500, n)) for n in (3, 0, 7)]
    its = [rng.random(len(m)) for m in mzs]
    p = PackedSpectra.from_arrays(mzs, its)
    for mzp, inp in ((0.0, 1.0), (0.5, 2.0)):
        a, b, norm = power_factors(p, mzp, inp)
        for x, (m, i) in enumerate(zip(mzs, its)):
            w = (m ** mzp) * (i ** inp)
            assert norm[x] == pytest.approx(np.sum(w ** 2) ** 0.5, rel=1e-15)
        assert (a is None) == (mzp == 0)
        assert np.array_equal(b, p.intensities ** inp)


def test_mgf_reader():
    from cudams.spectrum import load_from_mgf

    spectra = list(load_from_mgf(os.path.join(GOLDEN, "pesticides.mgf")))
    assert len(spectra) == 76
    n = [len(s.peaks) for s in spectra]
    assert min(n) == 16 and max(n) == 247 and sum(n) == 4721
    assert all(np.all(np.diff(s.peaks.mz) >= 0) for s in spectra)
    assert spectra[0].get("scans") == "675"
    assert spectra[0].get("SMILES") == "C1=CC=C2C(=C1)NC(=N2)C3=CC=CO3"


def test_uniform_generator():
    from cudams import synthetic

    p = synthetic.uniform_spectra(256, 1)
    n = p.lengths
    assert n.min() >= 10 and n.max() <= 200
    assert np.all(p.mz >= 50) and np.all(p.mz <= 1000)
    assert np.array_equal(p.mz.astype(np.float32).astype(np.float64), p.mz)
    mx = np.maximum.reduceat(p.intensities, p.offsets[:-1])
    assert np.all(mx == 1.0) and np.all(p.intensities > 0)
    q = synthetic.uniform_spectra(256, 1)
    assert np.array_equal(p.mz, q.mz)


def test_lognormal_generator_clusters():
    from cudams import synthetic

    p = synthetic.lognormal_spectra(2000, 5)
    n = p.lengths
    assert n.min() >= 5 and n.max() <= 1024
    assert 50 < np.median(n) < 80
    p.validate()
    # spectra with a cluster hold >= 30 peaks within 0.05 Da
    dense = 0
    for x in range(200):
        m = p.mz[p.offsets[x]:p.offsets[x + 1]]
        if len(m) >= 30 and np.min(m[29:] - m[:-29]) <= 0.05:
            dense += 1
    assert 15 <= dense <= 70


def test_fingerprint_generator_and_packing():
    from cudams import synthetic
    from cudams.similarity.CudaFingerprintSimilarity import pack_fingerprints

    fp = synthetic.fingerprints(64, 3, n_bits=2048, zero_rows=2)
    d = fp.mean(axis=1)
    assert np.all(d[:2] == 0) and 0.01 < d[2:].min() and d.max() < 0.35
    w = synthetic.pack_bits(fp)
    assert w.shape == (64, 64) and w.dtype == np.uint32
    for x in (2, 10):
        for b in (0, 31, 32, 1000, 2047):
            assert ((w[x, b // 32] >> (b % 32)) & 1) == fp[x, b]
    w2 = pack_fingerprints(list(fp), 2048)
    assert np.array_equal(w, w2)
    assert np.array_equal(synthetic.pack_bits(fp[:, :1000]), pack_fingerprints(list(fp[:, :1000]), 1000))


def test_sparse_container():
    from cudams.sparse import SparseScores

    sp = SparseScores(3, 4)
    data = np.rec.fromarrays([np.array([0.5, 0.9], np.float32), np.array([2, 3], np.int32)],
                             names=["score", "matches"])
    sp.add_sparse_data([0, 2], [1, 3], data, name="sparse")
    assert sp.score_names == ("sparse_score", "sparse_matches")
    a = sp.to_array("sparse_score")
    assert a.shape == (3, 4) and a[0, 1] == np.float32(0.5) and a[2, 3] == np.float32(0.9)
    with pytest.raises(ValueError):
        SparseScores(2, 2).add_sparse_data([5], [0], data[:1])


def test_argbatch():
    from cudams.utils import argbatch

    assert list(argbatch(list(range(5)), 2)) == [(0, 2), (2, 4), (4, 6)]


def test_workspace_size_depends_on_queries_only():
    """A prepared workspace must serve any later reference set: its size may
    not depend on the references (growing it would drop the prepared tiles)."""
    from cudams import _native

    L = _native.load()

    def ws(R, Q, peaks_q):
        r = _native.CsSpectra(R, R * 95, 8, 8, None, None, 8, 8, 1024, 1000.0)
        q = _native.CsSpectra(Q, peaks_q, 8, 8, None, None, 8, 8, 1024, 1000.0)
        n = ctypes.c_size_t()
        _native.check(L.cs_cosine_greedy_workspace_size(ctypes.byref(r), ctypes.byref(q),
                                                        ctypes.byref(n)))
        return n.value

    assert ws(17, 300, 28500) == ws(130, 300, 28500) == ws(500000, 300, 28500)
    assert ws(17, 600, 57000) > ws(17, 300, 28500)


def test_workspace_reuse_never_reallocates():
    import torch

    from cudams.device import Workspace

    w = Workspace("cpu")
    b = w.get(1000)
    assert w.get(500, keep=True) is b
    with pytest.raises(ValueError):
        w.get(5000, keep=True)


def _dummy_spectra(n, max_peaks):
    from cudams import _native

    fake = 0x1000  # never dereferenced: these calls fail in their host checks
    return _native.CsSpectra(n, n, fake, fake, None, None, fake, fake, max_peaks, 1000.0)


def test_lds_limits_rejected_before_any_launch():
    """Spectra of 65535 peaks with round_capacity 1024 need more LDS than the
    giant-pair kernel has: the call fails in its argument checks, before a
    kernel is queued (so no output is left half written)."""
    from cudams import _native

    L = _native.load()
    big, q = _dummy_spectra(4, 65535), _dummy_spectra(4, 65535)
    p = _native.CsCosineParams(0.1, 0.0, 2048, 0, 1024, 0, 0)
    st = L.cs_cosine_greedy(ctypes.byref(big), ctypes.byref(q), ctypes.byref(p), 0x1000, 0x1000,
                            0x1000, 4, None, 0, None)
    assert st == _native.CS_ERR_UNSUPPORTED
    assert b"LDS" in L.cs_last_error()


def test_batch_rejects_reference_blocks_beyond_24_bits():
    """A work item carries its reference block in 24 bits: a job that would
    need 2^24 or more blocks is refused at planning time."""
    from cudams import _native

    L = _native.load()
    job = _native.CsCosineJob()
    job.refs = _dummy_spectra((1 << 24) + 1, 8)
    job.queries = _dummy_spectra(1, 8)
    job.tolerance, job.shift, job.match_limit = 0.1, 0.0, 2048
    job.score = job.matches = job.overflow = 0x1000
    job.ld = 1
    bp = _native.CsBatchParams(0, 0, 1, 0)
    handle, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
    st = L.cs_cosine_batch_create(ctypes.byref(job), 1, ctypes.byref(bp), ctypes.byref(handle),
                                  ctypes.byref(nbytes))
    assert st == _native.CS_ERR_UNSUPPORTED
    assert b"reference blocks" in L.cs_last_error()
    assert not handle.value
