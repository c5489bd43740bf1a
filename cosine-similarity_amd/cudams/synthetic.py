"""Seeded synthetic workloads (numpy ``default_rng``), generated as CSR directly.

* :func:`uniform_spectra`    -- 10..200 peaks, m/z U[50, 1000] float32 sorted,
                                intensities U(0, 1] float32 scaled to max 1
* :func:`lognormal_spectra`  -- peaks round(lognormal(ln 60, 0.8)) clipped to
                                [5, 1024]; a fraction of spectra also carry a
                                dense cluster of 30..80 peaks inside 0.05 Da
* :func:`fingerprints`       -- binary rows, per-row density U[0.02, 0.3]
* :func:`batched_jobs`       -- 64 independent matrices, sizes log-uniform
* :func:`library_and_requests` -- a device library and a Poisson request stream
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .spectra import PackedSpectra


def _finish(rng, lens: np.ndarray, mz: np.ndarray) -> PackedSpectra:
    """Sort m/z within each spectrum, draw intensities, scale to max 1."""
    n = len(lens)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    seg = np.repeat(np.arange(n), lens)
    mz = mz.astype(np.float32)
    order = np.lexsort((mz, seg))
    mz = mz[order]
    it = (1.0 - rng.random(int(off[-1]))).astype(np.float32)  # U(0, 1]
    mx = np.zeros(n, np.float32)
    np.maximum.at(mx, seg, it)
    it = (it / mx[seg]).astype(np.float32)  # x / max(x) is exactly 1 at the max
    return PackedSpectra(mz.astype(np.float64), it.astype(np.float64), off)


def uniform_spectra(n: int, seed: int, n_min: int = 10, n_max: int = 200,
                    mz_lo: float = 50.0, mz_hi: float = 1000.0) -> PackedSpectra:
    rng = np.random.default_rng(seed)
    lens = rng.integers(n_min, n_max + 1, size=n).astype(np.int64)
    mz = rng.uniform(mz_lo, mz_hi, size=int(lens.sum()))
    return _finish(rng, lens, mz)


def lognormal_lengths(rng, n: int, median: float = 60.0, sigma: float = 0.8,
                      lo: int = 5, hi: int = 1024) -> np.ndarray:
    return np.clip(np.round(rng.lognormal(math.log(median), sigma, size=n)), lo, hi).astype(np.int64)


def lognormal_spectra(n: int, seed: int, cluster_fraction: float = 0.2,
                      cluster_peaks: Tuple[int, int] = (30, 80), cluster_span: float = 0.05,
                      mz_lo: float = 50.0, mz_hi: float = 1000.0, max_peaks: int = 1024,
                      median: float = 60.0, sigma: float = 0.8) -> PackedSpectra:
    rng = np.random.default_rng(seed)
    lens = lognormal_lengths(rng, n, median, sigma, 5, max_peaks)
    has_cluster = rng.random(n) < cluster_fraction
    k = np.where(has_cluster, rng.integers(cluster_peaks[0], cluster_peaks[1] + 1, size=n), 0)
    centre = rng.uniform(mz_lo, mz_hi, size=n)
    regular = np.minimum(lens, max_peaks - k)  # re-clip the total to max_peaks
    total = regular + k
    seg_reg = np.repeat(np.arange(n), regular)
    seg_cl = np.repeat(np.arange(n), k)
    mz_reg = rng.uniform(mz_lo, mz_hi, size=int(regular.sum()))
    mz_cl = centre[seg_cl] + rng.uniform(0.0, cluster_span, size=int(k.sum()))
    # interleave per spectrum: concatenate and let _finish sort within segments
    seg = np.concatenate([seg_reg, seg_cl])
    mz = np.concatenate([mz_reg, mz_cl])
    order = np.argsort(seg, kind="stable")
    return _finish(rng, total.astype(np.int64), mz[order])


def fingerprints(n: int, seed: int, n_bits: int = 2048, density: Tuple[float, float] = (0.02, 0.3),
                 zero_rows: int = 0) -> np.ndarray:
    """uint8 0/1 [n, n_bits]; the first ``zero_rows`` rows are all zero."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(density[0], density[1], size=(n, 1))
    fp = (rng.random((n, n_bits)) < p).astype(np.uint8)
    fp[:zero_rows] = 0
    return fp


def pack_bits(fp: np.ndarray) -> np.ndarray:
    """0/1 [n, n_bits] -> uint32 words [n, ceil(n_bits / 32)], bit b in word b // 32."""
    n, nb = fp.shape
    nw = (nb + 31) // 32
    out = np.zeros((n, nw * 4), np.uint8)
    packed = np.packbits(fp != 0, axis=1, bitorder="little")
    out[:, : packed.shape[1]] = packed
    return out.view(np.uint32)


@dataclass
class Job:
    """One matrix of a batch (the fields cudams.batch.CosineJob has)."""

    refs: PackedSpectra
    queries: PackedSpectra
    tolerance: float
    mz_power: float
    intensity_power: float
    shift: float = 0.0
    match_limit: int = 2048


def batched_jobs(seed: int = 6, n_jobs: int = 64, r_range=(16, 2048), q_range=(16, 4096),
                 tolerances=(0.01, 0.1), mz_power: float = 0.5) -> List[Job]:
    """Independent CosineGreedy matrices, R and Q log-uniform, peaks lognormal
    (no clusters), tolerance drawn per job."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        R = int(round(math.exp(rng.uniform(math.log(r_range[0]), math.log(r_range[1])))))
        Q = int(round(math.exp(rng.uniform(math.log(q_range[0]), math.log(q_range[1])))))
        tol = float(tolerances[int(rng.integers(len(tolerances)))])
        s = int(rng.integers(1 << 30))
        jobs.append(Job(lognormal_spectra(R, s, cluster_fraction=0.0),
                        lognormal_spectra(Q, s + 1, cluster_fraction=0.0), tol, mz_power, 1.0))
    return jobs


def library_and_requests(seed: int = 7, n_library: int = 500_000, rate: float = 200.0,
                         duration: float = 30.0, max_request: int = 256):
    """Library spectra (lognormal, no clusters) and a Poisson request stream:
    arrival times (s) and request sizes log-uniform on [1, max_request]."""
    rng = np.random.default_rng(seed)
    lib = lognormal_spectra(n_library, seed, cluster_fraction=0.0)
    n_req = int(rng.poisson(rate * duration))
    arrivals = np.sort(rng.uniform(0.0, duration, size=n_req))
    sizes = np.floor(np.exp(rng.uniform(0.0, math.log(max_request + 1), size=n_req))).astype(int)
    sizes = np.clip(sizes, 1, max_request)
    return lib, arrivals, sizes
