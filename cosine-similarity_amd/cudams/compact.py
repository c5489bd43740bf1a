"""Threshold compaction of a dense score block into COO arrays (host numpy).

Used by ``array_type="sparse"`` (reference CudaCosineGreedy.py:186-200): keeps
pairs with ``score >= threshold``; row indices are offset by ``row0``.
"""
from __future__ import annotations

import numpy as np
import torch


def compact_threshold(score: torch.Tensor, matches: torch.Tensor, overflow: torch.Tensor,
                      threshold: float, row0: int = 0):
    mask = score >= threshold
    r, q = torch.nonzero(mask, as_tuple=True)
    return ((r + row0).cpu().numpy().astype(np.int64), q.cpu().numpy().astype(np.int64),
            score[r, q].cpu().numpy(), matches[r, q].cpu().numpy(),
            overflow[r, q].cpu().numpy())
