"""Fingerprint similarity on the MI355X: drop-in for
cudams.similarity.CudaFingerprintSimilarity (reference
cudams/similarity/CudaFingerprintSimilarity.py:13-186).

Fingerprints are bit-packed on the host (one bit per position, nonzero = 1)
and compared on the matrix cores: the packed words feed fp4 (e2m1) MFMA
operands directly, one bit plane per scale, and the float32 accumulators hold
exact intersection counts.
"""
from __future__ import annotations

import warnings
from typing import List, Union

import numpy as np

from ..device import Workspace, fingerprint_similarity, require_gpu
from ..sparse import new_sparse
from ._base import BaseSimilarity


def pack_fingerprints(fps: List[np.ndarray], n_bits: int) -> np.ndarray:
    """0/1 rows -> uint32 words [rows, ceil(n_bits / 32)], bit b in word b // 32."""
    n_words = max(1, (n_bits + 31) // 32)
    out = np.zeros((len(fps), n_words * 4), np.uint8)
    nonbinary = False
    for x, fp in enumerate(fps):
        if fp is None:
            continue
        a = np.asarray(fp)
        if a.shape != (n_bits,):
            raise ValueError("fingerprints differ in length")
        bits = a != 0
        if not nonbinary and a.dtype != bool and np.any((a != 0) & (a != 1)):
            nonbinary = True
        packed = np.packbits(bits, bitorder="little")
        out[x, : len(packed)] = packed
    if nonbinary:
        warnings.warn("fingerprints hold values other than 0/1; nonzero entries count as set bits")
    return out.view(np.uint32)


class CudaFingerprintSimilarity(BaseSimilarity):
    """Jaccard / Dice / cosine between molecular fingerprints
    (``spectrum.get("fingerprint")``) of all (reference, query) pairs."""

    is_commutative = True
    score_datatype = np.float32

    def __init__(self, similarity_measure: str = "jaccard",
                 set_empty_scores: Union[float, int, str] = "nan"):
        self.set_empty_scores = set_empty_scores
        assert similarity_measure in ["cosine", "dice", "jaccard"], "Unknown similarity measure."
        self.similarity_measure = similarity_measure
        self._workspace = None

    def _empty_value(self) -> float:
        if self.set_empty_scores == "nan":
            return float("nan")
        if isinstance(self.set_empty_scores, (float, int)):
            return float(self.set_empty_scores)
        return 0.0  # the reference's create_full_matrix leaves zeros

    def _compute(self, fps_r, fps_q, valid_r, valid_q) -> np.ndarray:
        require_gpu()
        import torch

        lens = {len(np.asarray(f)) for f in fps_r + fps_q if f is not None}
        if len(lens) != 1:
            raise ValueError("fingerprints differ in length")
        n_bits = lens.pop()
        wr = torch.from_numpy(pack_fingerprints(fps_r, n_bits).view(np.int32)).cuda()
        wq = torch.from_numpy(pack_fingerprints(fps_q, n_bits).view(np.int32)).cuda()
        vr = torch.from_numpy(np.asarray(valid_r, np.uint8)).cuda()
        vq = torch.from_numpy(np.asarray(valid_q, np.uint8)).cuda()
        out = torch.empty((len(fps_r), len(fps_q)), dtype=torch.float32, device="cuda")
        if self._workspace is None:
            self._workspace = Workspace("cuda")
        fingerprint_similarity(self.similarity_measure, wr, wq, out, vr, vq, self._empty_value(),
                               workspace=self._workspace)
        return out.cpu().numpy()

    def pair(self, reference, query) -> np.ndarray:
        """1 x 1 float32 array, like the reference's pair() (:88-111)."""
        fr, fq = reference.get("fingerprint"), query.get("fingerprint")
        return self._compute([fr], [fq], [fr is not None], [fq is not None])

    def matrix(self, references: List, queries: List, array_type: str = "numpy",
               is_symmetric: bool = False):
        """Similarity of every (reference, query) pair, float32 [R, Q]; pairs
        with a missing fingerprint get ``set_empty_scores``."""
        fps_r = [s.get("fingerprint") for s in references]
        fps_q = [s.get("fingerprint") for s in queries]
        valid_r = [f is not None for f in fps_r]
        valid_q = [f is not None for f in fps_q]
        assert any(valid_r) and any(valid_q), (
            "Not enough molecular fingerprints.",
            "Apply 'add_fingerprint'filter first.",
        )
        if array_type not in ("numpy", "sparse"):
            raise NotImplementedError("Output array type is not yet implemented.")
        similarity_matrix = self._compute(fps_r, fps_q, valid_r, valid_q)
        if array_type == "sparse":
            scores_array = new_sparse(len(references), len(queries))
            scores_array.add_dense_matrix(similarity_matrix.astype(self.score_datatype), "")
            return scores_array
        return similarity_matrix.astype(self.score_datatype)
