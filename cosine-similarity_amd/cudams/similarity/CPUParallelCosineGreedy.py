"""CosineGreedy on the host cores: drop-in for
cudams.similarity.CPUParallelCosineGreedy.

Same constructor, ``score_datatype`` (score, matches; no overflow field) and
``pair``/``matrix`` signatures as the reference (cudams/similarity/
CPUParallelCosineGreedy.py:19-108).  The reference runs matchms's
collect_peak_pairs / score_best_matches per pair under numba's prange
(spectrum_similarity_functions.py:281-320); here the pairs run in the C++
engine libcsamd_cpu.so (OpenMP over query tiles x reference blocks), with the
tile peak index of the GPU path.  Results are matchms CosineGreedy's; no GPU
is needed.
"""
from __future__ import annotations

import ctypes
import warnings
from typing import List, Literal

import numpy as np

from .. import _native
from ..spectra import PackedSpectra, power_factors
from ._base import BaseSimilarity


class HostSpectra:
    """A PackedSpectra prepared for one (mz_power, intensity_power) in host
    memory, exposing the C struct ``cs_spectra`` with host pointers."""

    def __init__(self, packed: PackedSpectra, mz_power: float, intensity_power: float):
        mz_pow, int_pow, norm = power_factors(packed, mz_power, intensity_power)
        mz32 = packed.mz.astype(np.float32)
        exact = np.array_equal(mz32.astype(np.float64), packed.mz)

        def keep(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            return a if a.size else np.zeros(1, dtype=dtype)

        self.offsets = keep(packed.offsets, np.int64)
        self.mz = keep(mz32, np.float32)
        self.mz_exact = None if exact else keep(packed.mz, np.float64)
        self.mz_pow = None if mz_pow is None else keep(mz_pow, np.float64)
        self.int_pow = keep(int_pow, np.float64)
        self.norm = keep(norm, np.float64)
        self.n_spectra = packed.n_spectra
        self.n_peaks = int(packed.offsets[-1])
        self.max_peaks = int(packed.lengths.max()) if self.n_spectra else 0
        m = np.float32(np.abs(packed.mz).max()) if self.n_peaks else np.float32(0.0)
        if self.n_peaks and float(m) < float(np.abs(packed.mz).max()):
            m = np.nextafter(m, np.float32(np.inf))
        self.max_abs_mz = float(m)

    def struct(self) -> _native.CsSpectra:
        def ptr(a):
            return None if a is None else a.ctypes.data

        return _native.CsSpectra(self.n_spectra, self.n_peaks, ptr(self.offsets), ptr(self.mz),
                                 ptr(self.mz_exact), ptr(self.mz_pow), ptr(self.int_pow),
                                 ptr(self.norm), self.max_peaks, self.max_abs_mz)


class CPUParallelCosineGreedy(BaseSimilarity):
    """Greedy cosine score with peak m/z tolerance, all pairs on the CPU.

    ``match_limit`` and ``batch_size`` are kept for the reference's signature;
    as in the reference, neither changes a result (matchms collects every
    candidate).  ``n_threads`` (0 = OpenMP's default) is this
    implementation's own knob.
    """

    score_datatype = [("score", np.float32), ("matches", np.int32)]

    def __init__(
        self,
        tolerance: float = 0.1,
        mz_power: float = 0.0,
        intensity_power: float = 1.0,
        batch_size: int = 100,
        shift: float = 0,
        match_limit: int = 1024,
        verbose=False,
        n_threads: int = 0,
    ):
        self.tolerance = tolerance
        self.mz_power = mz_power
        self.int_power = intensity_power
        self.shift = shift
        self.match_limit = match_limit
        self.verbose = verbose
        if batch_size > 100:
            warnings.warn("Batch size larger than 100 might cause issues with numba.")
        self.batch_size = batch_size
        self.n_threads = n_threads
        self.kernel = _native.load_cpu()

    def _matrix_with_numpy_output(self, references: List, queries: List,
                                  is_symmetric: bool = False) -> np.ndarray:
        if is_symmetric:
            warnings.warn("is_symmetric is ignored here, it has no effect.")
        R = HostSpectra(PackedSpectra.from_spectra(references), self.mz_power, self.int_power)
        Q = R if queries is references else HostSpectra(PackedSpectra.from_spectra(queries),
                                                        self.mz_power, self.int_power)
        out = np.zeros((R.n_spectra, Q.n_spectra), dtype=self.score_datatype)
        if R.n_spectra and Q.n_spectra:
            score = np.zeros((R.n_spectra, Q.n_spectra), np.float32)
            matches = np.zeros((R.n_spectra, Q.n_spectra), np.int32)
            rs, qs = R.struct(), Q.struct()
            _native.check_cpu(self.kernel.cs_cpu_cosine_greedy(
                ctypes.byref(rs), ctypes.byref(qs), float(self.tolerance), float(self.shift),
                score.ctypes.data, matches.ctypes.data, Q.n_spectra, int(self.n_threads)))
            out["score"] = score
            out["matches"] = matches
        return out.view(np.recarray)

    def pair(self, reference, query) -> float:
        result_mat = self.matrix([reference], [query])
        return np.asarray(result_mat.squeeze(), dtype=self.score_datatype)

    def matrix(self, references: List, queries: List, array_type: Literal["numpy"] = "numpy",
               is_symmetric: bool = False, sparse_threshold: float = 0.75) -> np.ndarray:
        """Scores of every (reference, query) pair as an ``np.recarray``
        [R, Q] with fields score (f4) and matches (i4)."""
        assert self.kernel is not None, "Kernel isn't compiled - use .compile() first"
        if array_type == "numpy":
            return self._matrix_with_numpy_output(references, queries, is_symmetric=is_symmetric)
        raise NotImplementedError()


__all__ = ["CPUParallelCosineGreedy", "HostSpectra"]
