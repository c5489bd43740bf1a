"""CosineGreedy on the MI355X: drop-in for cudams.similarity.CudaCosineGreedy.

Same constructor, attributes, ``pair``/``matrix`` signatures and output dtypes
as the reference (cudams/similarity/CudaCosineGreedy.py:17-227); the compute
runs in libcsamd.so (gfx950 HIP), results follow matchms CosineGreedy.
"""
from __future__ import annotations

import warnings
from typing import List, Literal

import numpy as np

from ..device import Workspace, cosine_greedy, pack_records, require_gpu
from ..sparse import new_sparse
from ..spectra import DeviceSpectra, PackedSpectra
from ..utils import CudaTimer
from ._base import BaseSimilarity


class CudaCosineGreedy(BaseSimilarity):
    """Greedy cosine score with peak m/z tolerance, all pairs on the GPU.

    Parameters keep the reference's names and defaults (CudaCosineGreedy.py:
    24-35).  Differences in behaviour, all towards matchms:

    * ``matches`` and ``score`` are exact for every pair: a pair with more than
      ``match_limit`` candidate peak pairs is still computed in full.
      ``overflow`` keeps its meaning as "the pair had at least
      ``match_limit`` candidates" so code that inspects it sees the same
      pairs flagged.
    * ``batch_size`` only bounds the rows one device call produces for
      ``array_type="sparse"``; the dense matrix is produced by one call.
    """

    score_datatype = [
        ("score", np.float32),
        ("matches", np.int32),
        ("overflow", np.uint8),
    ]

    def __init__(
        self,
        tolerance: float = 0.1,
        mz_power: float = 0.0,
        intensity_power: float = 1.0,
        shift: float = 0,
        batch_size: int = 2048,
        n_max_peaks: int = 1024,
        match_limit: int = 2048,
        sparse_threshold: float = 0.75,
        verbose=False,
        light_candidates: int = 0,
        round_capacity: int = 0,
    ):
        self.tolerance = tolerance
        self.mz_power = mz_power
        self.int_power = intensity_power
        self.shift = shift
        self.batch_size = batch_size
        self.match_limit = match_limit
        self.verbose = verbose
        self.n_max_peaks = n_max_peaks
        self.kernel_time = 0  # seconds of device time of the last matrix() call
        self.sparse_threshold = sparse_threshold
        self.light_candidates = light_candidates
        self.round_capacity = round_capacity
        import torch

        self.device = "cuda" if torch.cuda.is_available() else "cpu"
        self._workspace = None
        if self.device != "cuda":
            warnings.warn(f"{self.__class__}: no ROCm GPU is visible.")

    def __str__(self) -> str:
        return (f"{self.__class__.__name__}(tolerance={self.tolerance}, mz_power={self.mz_power}, "
                f"intensity_power={self.int_power}, shift={self.shift}, "
                f"batch_size={self.batch_size}, n_max_peaks={self.n_max_peaks}, "
                f"match_limit={self.match_limit}, sparse_threshold={self.sparse_threshold})")

    # -- packing -------------------------------------------------------------
    def _to_device(self, spectra) -> DeviceSpectra:
        packed = PackedSpectra.from_spectra(spectra, self.n_max_peaks)
        return DeviceSpectra(packed, self.mz_power, self.int_power, self.device)

    def pair(self, reference, query) -> np.ndarray:
        result_mat = self.matrix([reference], [query])
        return np.asarray(result_mat.squeeze(), dtype=self.score_datatype)

    def matrix(
        self,
        references: List,
        queries: List,
        array_type: Literal["numpy", "sparse"] = "numpy",
        is_symmetric: bool = False,
    ):
        """Scores of every (reference, query) pair.

        ``array_type="numpy"`` returns an ``np.recarray`` [R, Q] with fields
        score (f4), matches (i4), overflow (u1).  ``"sparse"`` returns a COO
        container holding only pairs with ``score >= sparse_threshold``, fields
        named ``sparse_score`` / ``sparse_matches`` / ``sparse_overflow``.
        """
        if is_symmetric:
            warnings.warn("is_symmetric is ignored here, it has no effect.")
        if array_type == "sparse":
            assert self.sparse_threshold is not None, (
                "When using array_type `sparse` you have to set `score_threshold`. "
                "Scores below this will get discarded.")
            assert 0 <= self.sparse_threshold <= 1, "Score threshold must be in [0, 1] range."
        elif array_type != "numpy":
            raise NotImplementedError(f"array_type {array_type!r} is not implemented")
        require_gpu()
        import torch

        refs = self._to_device(references)
        ques = refs if queries is references else self._to_device(queries)
        R, Q = refs.n_spectra, ques.n_spectra
        if self._workspace is None:
            self._workspace = Workspace(self.device)
        timer = CudaTimer()
        if array_type == "numpy":
            score = torch.empty((R, Q), dtype=torch.float32, device=self.device)
            matches = torch.empty((R, Q), dtype=torch.int32, device=self.device)
            overflow = torch.empty((R, Q), dtype=torch.uint8, device=self.device)
            timer.start()
            cosine_greedy(refs, ques, self.tolerance, self.shift, self.match_limit, score, matches,
                          overflow, self._workspace, self.light_candidates, self.round_capacity)
            timer.stop()
            self.kernel_time = timer.elapsed_seconds()
            return self._download_records(score, matches, overflow)
        return self._sparse(refs, ques, timer)

    def _download_records(self, score, matches, overflow) -> np.ndarray:
        """The three device planes as one np.recarray: interleaved into packed
        9-byte records on the device, then one copy into pinned host memory
        that the returned array keeps (no host-side interleave)."""
        import torch

        R, Q = score.shape
        dtype = np.dtype(self.score_datatype)
        assert dtype.itemsize == 9
        if R * Q == 0:
            return np.zeros((R, Q), dtype=dtype).view(np.recarray)
        records = torch.empty(9 * R * Q, dtype=torch.uint8, device=score.device)
        pack_records(score, matches, overflow, records)
        host = torch.empty(9 * R * Q, dtype=torch.uint8, pin_memory=True)
        host.copy_(records, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return host.numpy().view(dtype).reshape(R, Q).view(np.recarray)

    def _sparse(self, refs: DeviceSpectra, ques: DeviceSpectra, timer):
        import torch

        from ..device import threshold_compact

        R, Q = refs.n_spectra, ques.n_spectra
        rows = max(1, min(R, int(self.batch_size)))
        parts = []
        self.kernel_time = 0.0
        score = torch.empty((rows, Q), dtype=torch.float32, device=self.device)
        matches = torch.empty((rows, Q), dtype=torch.int32, device=self.device)
        overflow = torch.empty((rows, Q), dtype=torch.uint8, device=self.device)
        for r0 in range(0, R, rows):
            r1 = min(R, r0 + rows)
            sub = refs.slice(r0, r1)
            timer.start()
            cosine_greedy(sub, ques, self.tolerance, self.shift, self.match_limit, score, matches,
                          overflow, self._workspace, self.light_candidates, self.round_capacity)
            timer.stop()
            self.kernel_time += timer.elapsed_seconds()
            parts.append(threshold_compact(score[: r1 - r0], matches[: r1 - r0],
                                           overflow[: r1 - r0], self.sparse_threshold, r0))
        r = np.concatenate([p[0] for p in parts])
        q = np.concatenate([p[1] for p in parts])
        s = np.concatenate([p[2] for p in parts])
        m = np.concatenate([p[3] for p in parts])
        o = np.concatenate([p[4] for p in parts])
        sp = new_sparse(R, Q)
        sp.add_sparse_data(r, q, np.rec.fromarrays([s, m, o], names=["score", "matches",
                                                                     "overflow"]), name="sparse")
        return sp
