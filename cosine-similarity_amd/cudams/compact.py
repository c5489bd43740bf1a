"""Threshold compaction of a dense score block into COO arrays.

Used by ``array_type="sparse"`` (reference CudaCosineGreedy.py:186-200): keeps
pairs with ``score >= threshold`` in row-major order; row indices are offset
by ``row0``.  Runs on the device (cs_threshold_compact); only the kept entries
travel to the host.
"""
from __future__ import annotations

import torch

from .device import threshold_compact


def compact_threshold(score: torch.Tensor, matches: torch.Tensor, overflow: torch.Tensor,
                      threshold: float, row0: int = 0):
    return threshold_compact(score, matches, overflow, threshold, row0)
