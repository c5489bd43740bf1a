"""Batches of independent CosineGreedy matrices in one persistent launch.

The reference computes one matrix per ``CudaCosineGreedy.matrix`` call, with
one kernel launch and one host sync per tile pair
(cudams/similarity/CudaCosineGreedy.py:149-185).  Many small, uneven matrices
then cost mostly launch latency and idle compute units.  Here:

* :class:`BatchRunner` plans all jobs together (``cs_cosine_batch_create``):
  work items are (job, 64-query tile, block of references), costliest first;
  one persistent lane-kernel launch takes them from a shared counter, and one
  launch each settles the queued many-candidate pairs of every job.  The
  tables are uploaded once; :meth:`BatchRunner.run` only launches kernels, so
  it is captured into a HIP graph (:meth:`BatchRunner.capture`) and replayed.
* :func:`per_job` is the baseline: one ``cs_cosine_greedy`` call per job.
* :func:`autotune` times candidate launch shapes (register candidates per
  lane, giant-pair round capacity, references per work item) on the batch
  itself, with graph replays, and keeps the fastest per workload signature.
"""
from __future__ import annotations

import ctypes
import dataclasses
import itertools
import json
import os
import time
from dataclasses import dataclass
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from .device import Workspace, cosine_greedy, require_gpu
from .spectra import DeviceSpectra, PackedSpectra


@dataclass
class CosineJob:
    """One independent matrix: references x queries at its own tolerance."""

    refs: PackedSpectra
    queries: PackedSpectra
    tolerance: float = 0.1
    shift: float = 0.0
    mz_power: float = 0.0
    intensity_power: float = 1.0
    match_limit: int = 2048


@dataclass(frozen=True)
class BatchParams:
    """cs_batch_params; 0 = library default."""

    light_candidates: int = 0
    round_capacity: int = 0
    refs_per_block: int = 0
    grid: int = 0

    def struct(self):
        return _native.CsBatchParams(self.light_candidates, self.round_capacity,
                                     self.refs_per_block, self.grid)


class _DeviceJob:
    def __init__(self, job: CosineJob, device):
        self.job = job
        self.R = DeviceSpectra(job.refs, job.mz_power, job.intensity_power, device)
        self.Q = self.R if job.queries is job.refs else \
            DeviceSpectra(job.queries, job.mz_power, job.intensity_power, device)
        n, m = self.R.n_spectra, self.Q.n_spectra
        self.score = torch.empty((max(n, 1), max(m, 1)), dtype=torch.float32, device=device)
        self.matches = torch.empty((max(n, 1), max(m, 1)), dtype=torch.int32, device=device)
        self.overflow = torch.empty((max(n, 1), max(m, 1)), dtype=torch.uint8, device=device)

    def struct(self):
        j = self.job
        return _native.CsCosineJob(self.R.struct(), self.Q.struct(), float(j.tolerance),
                                   float(j.shift), int(j.match_limit), self.score.data_ptr(),
                                   self.matches.data_ptr(), self.overflow.data_ptr(),
                                   self.score.stride(0))

    def result(self):
        n, m = self.R.n_spectra, self.Q.n_spectra
        return (self.score[:n, :m].cpu().numpy(), self.matches[:n, :m].cpu().numpy(),
                self.overflow[:n, :m].cpu().numpy())


class BatchRunner:
    """All jobs planned, uploaded and computed together (see module doc)."""

    def __init__(self, jobs: Sequence[CosineJob], params: BatchParams = BatchParams(),
                 device="cuda", _device_jobs: Optional[List[_DeviceJob]] = None):
        require_gpu()
        self.device = torch.device(device)
        self.params = params
        self.djobs = _device_jobs or [_DeviceJob(j, self.device) for j in jobs]
        L = _native.load()
        arr = (_native.CsCosineJob * len(self.djobs))(*[d.struct() for d in self.djobs])
        plan = ctypes.c_void_p()
        nbytes = ctypes.c_size_t()
        bp = params.struct()
        _native.check(L.cs_cosine_batch_create(arr, len(self.djobs), ctypes.byref(bp),
                                               ctypes.byref(plan), ctypes.byref(nbytes)))
        self._plan = plan
        self.workspace = torch.empty(max(nbytes.value, 256), dtype=torch.uint8,
                                     device=self.device)
        s = torch.cuda.current_stream(self.device)
        _native.check(L.cs_cosine_batch_upload(plan, self.workspace.data_ptr(),
                                               self.workspace.numel(), s.cuda_stream))
        self.graph: Optional[torch.cuda.CUDAGraph] = None

    def run(self, stream=None) -> None:
        s = stream or torch.cuda.current_stream(self.device)
        _native.check(_native.load().cs_cosine_batch_run(self._plan, s.cuda_stream))

    def capture(self) -> "torch.cuda.CUDAGraph":
        """Capture one run into a graph (after one eager run, which sets up
        every kernel attribute outside the capture)."""
        self.run()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self) -> None:
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def results(self) -> List[Tuple[np.ndarray, np.ndarray, np.ndarray]]:
        torch.cuda.synchronize(self.device)
        return [d.result() for d in self.djobs]

    def close(self) -> None:
        if self._plan:
            _native.load().cs_cosine_batch_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def per_job(djobs: Sequence[_DeviceJob], workspace: Optional[Workspace] = None,
            params: BatchParams = BatchParams()) -> None:
    """The baseline: one cs_cosine_greedy call (its own launches) per job, in order."""
    ws = workspace or Workspace(djobs[0].score.device if djobs else "cuda")
    for d in djobs:
        if d.R.n_spectra == 0 or d.Q.n_spectra == 0:
            continue
        j = d.job
        cosine_greedy(d.R, d.Q, j.tolerance, j.shift, j.match_limit, d.score, d.matches,
                      d.overflow, ws, params.light_candidates, params.round_capacity,
                      refs_per_block=params.refs_per_block)


def signature(jobs: Sequence[CosineJob]) -> str:
    """A coarse description of a batch: job count, total pairs and peaks,
    tolerances; batches with the same signature share a tuned shape."""
    pairs = sum(j.refs.n_spectra * j.queries.n_spectra for j in jobs)
    peaks = sum(int(j.refs.offsets[-1]) + int(j.queries.offsets[-1]) for j in jobs)
    tols = sorted({round(float(j.tolerance), 6) for j in jobs})
    return f"jobs={len(jobs)};pairs={pairs:.3g};peaks={peaks:.3g};tol={tols}"


DEFAULT_SPACE = {
    "light_candidates": (4, 8),
    "refs_per_block": (64, 128, 256, 512, 1024, 2048),
    "round_capacity": (256,),
}


def time_runner(runner: BatchRunner, replays: int = 5) -> float:
    """Milliseconds per graph replay (HIP events on the replay stream)."""
    runner.capture()
    runner.replay()
    torch.cuda.synchronize(runner.device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(replays):
        runner.replay()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / replays


def autotune(jobs: Sequence[CosineJob], space: Dict[str, Iterable[int]] = None,
             replays: int = 3, cache_path: Optional[str] = None, device="cuda",
             verbose: bool = False) -> Tuple[BatchParams, Dict[BatchParams, float]]:
    """Grid search over ``space`` on this batch; returns the fastest
    BatchParams and every timing.  With ``cache_path`` (JSON), a shape found
    for the same device and signature is reused without timing."""
    key = f"{torch.cuda.get_device_name(device)}|{signature(jobs)}"
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if key in cache:
            return BatchParams(**cache[key]), {}
    space = space or DEFAULT_SPACE
    names = list(space)
    djobs = [_DeviceJob(j, torch.device(device)) for j in jobs]
    timings: Dict[BatchParams, float] = {}
    for values in itertools.product(*(space[n] for n in names)):
        bp = BatchParams(**dict(zip(names, values)))
        runner = BatchRunner(jobs, bp, device, _device_jobs=djobs)
        timings[bp] = time_runner(runner, replays)
        runner.close()
        if verbose:
            print(f"[autotune] {bp}: {timings[bp]:.3f} ms", flush=True)
    best = min(timings, key=timings.get)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                cache = json.load(fh)
        cache[key] = dataclasses.asdict(best)
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return best, timings
