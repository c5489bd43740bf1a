"""Base class of the similarity plug-ins.

When matchms is importable the plug-ins subclass its ``BaseSimilarity`` so
``matchms.calculate_scores`` accepts them unchanged (the reference does so at
CudaCosineGreedy.py:8,17 and CudaFingerprintSimilarity.py:5,13).  Otherwise a
stand-in with the same protocol (``pair``, ``matrix``, ``is_commutative``,
``score_datatype``) is used.
"""
from __future__ import annotations

import numpy as np

try:  # pragma: no cover - matchms is not installed in this image
    from matchms.similarity.BaseSimilarity import BaseSimilarity  # type: ignore
except Exception:  # noqa: BLE001
    class BaseSimilarity:  # type: ignore[no-redef]
        """Protocol of matchms BaseSimilarity used by the plug-ins."""

        is_commutative = True
        score_datatype = np.float64

        def pair(self, reference, query):
            raise NotImplementedError

        def matrix(self, references, queries, array_type="numpy", is_symmetric=False):
            raise NotImplementedError
