"""COO score container for ``array_type="sparse"`` results.

The reference returns ``sparsestack.StackedSparseArray`` (CudaCosineGreedy.py:
206-227, CudaFingerprintSimilarity.py:178-183).  sparsestack is used when it is
importable; otherwise :class:`SparseScores` offers the calls the reference and
matchms make on it: ``add_sparse_data``, ``add_dense_matrix``, ``to_array``,
``row``, ``col``, ``data``, ``score_names`` and ``shape``.
"""
from __future__ import annotations

import numpy as np


class SparseScores:
    def __init__(self, n_row: int, n_col: int):
        self.shape = (int(n_row), int(n_col))
        self.row = np.zeros(0, np.int64)
        self.col = np.zeros(0, np.int64)
        self.data = None

    @property
    def score_names(self):
        return () if self.data is None else tuple(self.data.dtype.names)

    def _rename(self, data: np.ndarray, name: str) -> np.ndarray:
        if data.dtype.names is None:
            return np.rec.fromarrays([data], names=[name or "score"])
        names = [f"{name}_{n}" if name else n for n in data.dtype.names]
        return np.rec.fromarrays([data[n] for n in data.dtype.names], names=names)

    def add_sparse_data(self, row, col, data, name: str = ""):
        if self.data is not None:
            raise ValueError("this container already holds a score layer")
        row = np.asarray(row, np.int64)
        col = np.asarray(col, np.int64)
        if row.shape != col.shape or len(row) != len(data):
            raise ValueError("row, col and data must have the same length")
        if len(row) and (row.min() < 0 or col.min() < 0 or row.max() >= self.shape[0]
                         or col.max() >= self.shape[1]):
            raise ValueError("index outside the matrix")
        self.row, self.col = row, col
        self.data = self._rename(np.asarray(data), name)
        return self

    def add_dense_matrix(self, matrix: np.ndarray, name: str = ""):
        m = np.asarray(matrix)
        r, c = np.nonzero(np.ones(m.shape[:2], bool))
        return self.add_sparse_data(r, c, m[r, c], name)

    def to_array(self, name: str = None) -> np.ndarray:
        if self.data is None:
            return np.zeros(self.shape)
        if name is None:
            if len(self.score_names) != 1:
                raise ValueError(f"name one of {self.score_names}")
            name = self.score_names[0]
        out = np.zeros(self.shape, dtype=self.data[name].dtype)
        out[self.row, self.col] = self.data[name]
        return out


def new_sparse(n_row: int, n_col: int):
    try:  # pragma: no cover - sparsestack is not installed in this image
        from sparsestack import StackedSparseArray  # type: ignore

        return StackedSparseArray(n_row, n_col)
    except Exception:  # noqa: BLE001
        return SparseScores(n_row, n_col)
