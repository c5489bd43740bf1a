"""Small helpers kept from the reference API (cudams/utils.py)."""
from __future__ import annotations

import time
from typing import Iterable, Tuple


def argbatch(lst, batch_size: int) -> Iterable[Tuple[int, int]]:
    """(start, end) index pairs of ``batch_size`` chunks of ``lst``
    (reference cudams/utils.py:24-35; ``end`` may exceed ``len(lst)``)."""
    for i in range(0, len(lst), batch_size):
        yield i, i + batch_size


class Timer:
    """Wall-clock context manager; ``duration`` in seconds (cudams/utils.py:267-274)."""

    def __enter__(self):
        self.duration = -time.perf_counter()
        return self

    def __exit__(self, *args):
        self.duration += time.perf_counter()


class CudaTimer:
    """Device timer from HIP events recorded on torch's current stream (the
    stream the plug-ins launch on); cudams/utils.py:322-362."""

    def __init__(self):
        import torch

        self._start = torch.cuda.Event(enable_timing=True)
        self._stop = torch.cuda.Event(enable_timing=True)

    def start(self):
        self._start.record()

    def stop(self):
        self._stop.record()

    def elapsed(self) -> float:
        """Milliseconds between the last start and stop (synchronises on stop)."""
        self._stop.synchronize()
        return self._start.elapsed_time(self._stop)

    def elapsed_seconds(self) -> float:
        return self.elapsed() / 1000.0
