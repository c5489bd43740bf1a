"""Served library search: a spectral library resident on the GPU, searched by
a stream of query requests with continuous batching.

The reference has no serving loop: its users call ``CudaCosineGreedy.matrix``
per request and take the best hits per query from the dense result
(tests/usecase/test_tutorial_pesticide.py:38-41).  Here the library is packed,
uploaded and indexed once (the query-tile index of cs_cosine_greedy_indexed,
with the library on the tiled side), and request spectra stream through it as
references:

* :class:`LibraryIndex` -- the device library and its prepared index;
* :class:`SearchServer` -- continuous batching.  A dispatcher thread takes a
  free *slot* (its own stream, call scratch, output block and pinned staging
  buffers, sized by a memory budget), fuses every queued request that fits
  into it, and launches search + top-k on the slot's stream; the slot's own
  completion thread hands each request its rows when the slot's event fires
  and frees the slot (a small batch is not held behind a larger one launched
  before it).  Requests that arrive while batches run are admitted into the
  next slot that frees up, and batches on different slots overlap on the GPU;
* :class:`PlainSearch` -- the baseline: one search + top-k call per request,
  one request at a time, on the same prepared index.

A request's hits do not depend on what it was fused with: every pair is
computed independently, and the top-k order is total (score descending, then
library index ascending).
"""
from __future__ import annotations

import gc
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native
from .device import (Workspace, cosine_greedy_indexed, cosine_greedy_indexed_bytes,
                     require_gpu, topk_rows)
from .spectra import DeviceSpectra, PackedSpectra, power_factors


@dataclass
class SearchHits:
    """Top-k library hits of each spectrum of one request (row = spectrum)."""

    scores: np.ndarray   # float32 [n, k], NaN where fewer than k library spectra
    indices: np.ndarray  # int32 [n, k], library index, -1 where empty
    matches: np.ndarray  # int32 [n, k]


@dataclass
class _Request:
    """Host-packed request spectra (CSR with local offsets)."""

    offsets: np.ndarray   # int64 [n + 1], offsets[0] == 0
    mz32: np.ndarray      # float32 [P]
    mz64: np.ndarray      # float64 [P]
    mz_pow: Optional[np.ndarray]
    int_pow: np.ndarray   # float64 [P]
    norm: np.ndarray      # float64 [n]
    exact: bool
    max_abs_mz: float
    max_peaks: int
    t_arrival: float = 0.0
    t_done: float = 0.0
    future: Future = field(default_factory=Future)

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def peaks(self) -> int:
        return int(self.offsets[-1])


def pack_request(spectra, mz_power: float = 0.0, intensity_power: float = 1.0,
                 n_max_peaks: Optional[int] = 1024) -> _Request:
    """Host packing of one request: a PackedSpectra or a sequence of
    spectrum-like objects (first ``n_max_peaks`` peaks of each)."""
    p = spectra if isinstance(spectra, PackedSpectra) else \
        PackedSpectra.from_spectra(spectra, n_max_peaks)
    mz_pow, int_pow, norm = power_factors(p, mz_power, intensity_power)
    mz32 = p.mz.astype(np.float32)
    return _Request(
        offsets=np.ascontiguousarray(p.offsets - p.offsets[0], np.int64),
        mz32=mz32, mz64=np.ascontiguousarray(p.mz, np.float64), mz_pow=mz_pow,
        int_pow=int_pow, norm=norm,
        exact=bool(np.array_equal(mz32.astype(np.float64), p.mz)),
        max_abs_mz=float(np.abs(p.mz).max()) if len(p.mz) else 0.0,
        max_peaks=int(p.lengths.max()) if p.n_spectra else 0)


def _f32_up(x: float) -> float:
    m = np.float32(x)
    if float(m) < x:
        m = np.nextafter(m, np.float32(np.inf))
    return float(m)


class LibraryIndex:
    """A library of spectra uploaded once, with its query-tile index prepared
    once (CS_STAGE_PREPARE of cs_cosine_greedy_indexed)."""

    def __init__(self, library: PackedSpectra, tolerance: float = 0.01, shift: float = 0.0,
                 mz_power: float = 0.0, intensity_power: float = 1.0, match_limit: int = 2048,
                 device="cuda"):
        require_gpu()
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.tolerance, self.shift = float(tolerance), float(shift)
        self.mz_power, self.intensity_power = float(mz_power), float(intensity_power)
        self.match_limit = int(match_limit)
        self.lib = DeviceSpectra(library, mz_power, intensity_power, self.device)
        self.n = self.lib.n_spectra
        ib, sb = cosine_greedy_indexed_bytes(self.lib)
        self.index = torch.empty(max(ib, 256), dtype=torch.uint8, device=self.device)
        self.scratch_bytes = max(sb, 256)
        empty = DeviceSpectra(PackedSpectra.from_arrays([], []), mz_power, intensity_power,
                              self.device)
        cosine_greedy_indexed(empty, self.lib, self.tolerance, self.shift, self.match_limit,
                              None, None, None, self.index, None,
                              stages=_native.CS_STAGE_PREPARE)
        torch.cuda.synchronize(self.device)

    def search(self, refs_struct, n_refs: int, score, matches, overflow, scratch,
               stream=None) -> None:
        """LIGHT | HEAVY of ``n_refs`` request spectra against the library."""
        L = _native.load()
        import ctypes

        qs = self.lib.struct()
        p = _native.CsCosineParams(self.tolerance, self.shift, self.match_limit, 0, 0, 0,
                                   _native.CS_STAGE_LIGHT | _native.CS_STAGE_HEAVY)
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        _native.check(L.cs_cosine_greedy_indexed(
            ctypes.byref(refs_struct), ctypes.byref(qs), ctypes.byref(p), score.data_ptr(),
            matches.data_ptr(), overflow.data_ptr(), score.stride(0), self.index.data_ptr(),
            self.index.numel(), scratch.data_ptr(), scratch.numel(), s))


def _a256(x: int) -> int:
    return (x + 255) & ~255


class _Slot:
    """Buffers of one in-flight fused batch.  A batch's request arrays are
    packed back to back into one pinned staging buffer and uploaded with one
    copy, and its top-k rows come back with one copy."""

    def __init__(self, ix: LibraryIndex, spectra: int, peaks: int, k: int):
        dev = ix.device
        self.ix, self.cap, self.peak_cap, self.k = ix, spectra, peaks, k
        self.stream = torch.cuda.Stream(dev)  # default priority
        self.scratch = torch.empty(ix.scratch_bytes, dtype=torch.uint8, device=dev)
        self.score = torch.empty((spectra, ix.n), dtype=torch.float32, device=dev)
        self.matches = torch.empty((spectra, ix.n), dtype=torch.int32, device=dev)
        self.overflow = torch.empty((spectra, ix.n), dtype=torch.uint8, device=dev)
        self.topk_ws = Workspace(dev)
        import ctypes

        nb = ctypes.c_size_t(0)  # sized once for a full slot: never grows in flight
        _native.check(_native.load().cs_topk_rows_workspace_size(spectra, ix.n, k,
                                                                  ctypes.byref(nb)))
        self.topk_ws.get(nb.value)
        # offsets, m/z (f32), int_pow, norms, exact m/z, m/z powers
        pk = max(peaks, 1)
        self.in_bytes = (_a256(8 * (spectra + 1)) + _a256(4 * pk) + 3 * _a256(8 * pk) +
                         _a256(8 * max(spectra, 1)))
        self.h_in = torch.empty(self.in_bytes, dtype=torch.uint8, pin_memory=True)
        self.d_in = torch.empty(self.in_bytes, dtype=torch.uint8, device=dev)
        # scores, library indices, matches: [3, rows, k] words of the batch's rows
        self.d_out = torch.empty(3 * spectra * k, dtype=torch.int32, device=dev)
        self.h_out = torch.empty(3 * spectra * k, dtype=torch.int32, pin_memory=True)
        self.done = torch.cuda.Event()
        self.batch: List[_Request] = []
        self.rows = 0

    def launch(self, batch: List[_Request]) -> None:
        """Fuse ``batch`` into this slot and enqueue upload, search, top-k and
        download on the slot's stream."""
        n = sum(r.n for r in batch)
        P = sum(r.peaks for r in batch)
        assert n <= self.cap and P <= self.peak_cap
        exact = all(r.exact for r in batch)
        mzp = self.ix.mz_power != 0.0
        # this batch's segments of the staging buffer, back to back
        seg, o = {}, 0
        for name, size, dt in (("off", 8 * (n + 1), np.int64), ("mz", 4 * P, np.float32),
                               ("inp", 8 * P, np.float64), ("norm", 8 * n, np.float64),
                               ("mzx", 0 if exact else 8 * P, np.float64),
                               ("mzp", 8 * P if mzp else 0, np.float64)):
            seg[name] = (o, size, dt)
            o += _a256(size)
        hb = self.h_in.numpy()
        h = {name: hb[a:a + size].view(dt) for name, (a, size, dt) in seg.items()}
        off, po, s = h["off"], 0, 0
        off[0] = 0
        for r in batch:
            off[s + 1:s + r.n + 1] = r.offsets[1:] + po
            h["mz"][po:po + r.peaks] = r.mz32
            if not exact:
                h["mzx"][po:po + r.peaks] = r.mz64
            if mzp:
                h["mzp"][po:po + r.peaks] = r.mz_pow
            h["inp"][po:po + r.peaks] = r.int_pow
            h["norm"][s:s + r.n] = r.norm
            po += r.peaks
            s += r.n
        self.batch, self.rows = batch, n
        st = self.stream
        base = self.d_in.data_ptr()

        def ptr(name):  # m/z powers and exact m/z are optional
            a, size, _ = seg[name]
            return base + a if size or name not in ("mzx", "mzp") else None

        k = self.k
        ts = self.d_out[:n * k].view(torch.float32).view(n, k)
        ti = self.d_out[n * k:2 * n * k].view(n, k)
        tm = self.d_out[2 * n * k:3 * n * k].view(n, k)
        with torch.cuda.stream(st):
            self.d_in[:o].copy_(self.h_in[:o], non_blocking=True)
            refs = _native.CsSpectra(
                n, P, ptr("off"), ptr("mz"), ptr("mzx"), ptr("mzp"), ptr("inp"), ptr("norm"),
                max((r.max_peaks for r in batch), default=0),
                _f32_up(max((r.max_abs_mz for r in batch), default=0.0)))
            self.ix.search(refs, n, self.score, self.matches, self.overflow, self.scratch, st)
            topk_rows(self.score, k, self.matches, rows=n, cols=self.ix.n, stream=st,
                      workspace=self.topk_ws, out=(ts, ti, tm))
            self.h_out[:3 * n * k].copy_(self.d_out[:3 * n * k], non_blocking=True)
            self.done.record(st)

    def finish(self) -> None:
        """Wait for the batch, then hand every request its rows."""
        self.done.synchronize()
        t = time.perf_counter()
        n, k = self.rows, self.k
        out = self.h_out.numpy()
        rs = out[:n * k].view(np.float32).reshape(n, k)
        ri = out[n * k:2 * n * k].reshape(n, k)
        rm = out[2 * n * k:3 * n * k].reshape(n, k)
        s = 0
        for r in self.batch:
            hits = SearchHits(rs[s:s + r.n].copy(), ri[s:s + r.n].copy(), rm[s:s + r.n].copy())
            s += r.n
            r.t_done = t
            r.future.set_result(hits)
        self.batch, self.rows = [], 0


class SearchServer:
    """Continuous-batching search over a :class:`LibraryIndex`.

    ``memory_budget`` (bytes of device memory for in-flight outputs) is split
    over ``slots``; a slot holds ``budget / slots / (9 * library size)``
    request spectra (score, matches and overflow of every library pair).  A
    request larger than one slot is rejected at submit.
    """

    def __init__(self, index: LibraryIndex, k: int = 10, slots: int = 2,
                 memory_budget: int = 6 << 30, max_batch_spectra: int = 512,
                 max_batch_peaks: Optional[int] = None):
        self.ix, self.k = index, int(k)
        per_spec = 9 * index.n
        cap = int(min(max_batch_spectra, memory_budget // max(slots, 1) // max(per_spec, 1)))
        if cap < 1:
            raise ValueError("memory budget holds no request spectrum")
        self.slot_spectra = cap
        self.slot_peaks = int(max_batch_peaks or cap * 1024)
        self.slots = [_Slot(index, cap, self.slot_peaks, self.k) for _ in range(slots)]
        self.free: "queue.Queue[_Slot]" = queue.Queue()
        for s in self.slots:
            self.free.put(s)
        self.pending: "queue.Queue[Optional[_Request]]" = queue.Queue()
        # one completion queue per slot: a slot is finished as soon as its own
        # batch is done, not behind a larger batch launched before it
        self.launched = {id(s): queue.Queue() for s in self.slots}
        self._carry: Optional[_Request] = None
        self._threads: List[threading.Thread] = []
        self.batches = 0
        self.batch_sizes: List[int] = []
        self.error: Optional[BaseException] = None

    # -- client side ---------------------------------------------------------
    def submit(self, request, t_arrival: Optional[float] = None) -> Future:
        """Queue a request (a packed request or spectra); its future resolves
        to :class:`SearchHits`."""
        if self.error is not None:
            raise RuntimeError("the search server failed") from self.error
        r = request if isinstance(request, _Request) else \
            pack_request(request, self.ix.mz_power, self.ix.intensity_power)
        if r.n > self.slot_spectra or r.peaks > self.slot_peaks:
            raise ValueError(f"request of {r.n} spectra / {r.peaks} peaks exceeds a slot "
                             f"({self.slot_spectra} / {self.slot_peaks})")
        r.future = Future()
        r.t_arrival = time.perf_counter() if t_arrival is None else t_arrival
        if r.n == 0:
            r.t_done = time.perf_counter()
            r.future.set_result(SearchHits(np.zeros((0, self.k), np.float32),
                                           np.zeros((0, self.k), np.int32),
                                           np.zeros((0, self.k), np.int32)))
            return r.future
        self.pending.put(r)
        return r.future

    # -- server threads ------------------------------------------------------
    def _take_batch(self) -> Optional[List[_Request]]:
        first = self._carry or self.pending.get()
        self._carry = None
        if first is None:
            return None
        batch, n, P = [first], first.n, first.peaks
        while True:
            try:
                r = self.pending.get_nowait()
            except queue.Empty:
                break
            if r is None:
                self.pending.put(None)
                break
            if n + r.n > self.slot_spectra or P + r.peaks > self.slot_peaks:
                self._carry = r  # first of the next batch
                break
            batch.append(r)
            n += r.n
            P += r.peaks
        return batch

    def _fail_pending(self, e: BaseException) -> None:
        reqs = [self._carry] if self._carry is not None else []
        self._carry = None
        while True:
            try:
                reqs.append(self.pending.get_nowait())
            except queue.Empty:
                break
        for r in reqs:
            if r is not None and not r.future.done():
                r.future.set_exception(e)

    def _dispatch(self) -> None:
        try:
            torch.cuda.set_device(self.ix.device)
            while True:
                slot = self.free.get()       # admitted only when a slot is free
                batch = self._take_batch()   # everything queued that fits
                if batch is None:
                    break
                slot.launch(batch)
                self.batches += 1
                self.batch_sizes.append(sum(r.n for r in batch))
                self.launched[id(slot)].put(slot)
        except BaseException as e:  # surface in the client threads
            self.error = e
            self._fail_pending(e)
        for q in self.launched.values():
            q.put(None)

    def _complete(self, q: "queue.Queue[Optional[_Slot]]") -> None:
        """Completion of one slot's batches: wait for its event, hand each
        request its rows, free the slot."""
        try:
            torch.cuda.set_device(self.ix.device)
        except BaseException as e:
            self.error = e
        while True:
            slot = q.get()
            if slot is None:
                break
            try:
                slot.finish()
            except BaseException as e:
                self.error = e
                for r in slot.batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            self.free.put(slot)

    def start(self) -> "SearchServer":
        workers = [(self._dispatch, ())] + [(self._complete, (q,)) for q in self.launched.values()]
        for fn, args in workers:
            t = threading.Thread(target=fn, args=args, daemon=True)
            t.start()
            self._threads.append(t)
        return self

    def stop(self) -> None:
        self.pending.put(None)
        for t in self._threads:
            t.join()
        self._threads = []
        if self.error is not None:
            raise self.error

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()


class PlainSearch:
    """Baseline: one search + top-k call per request, synchronous, on the
    same prepared library index."""

    def __init__(self, index: LibraryIndex, k: int = 10):
        self.ix, self.k = index, int(k)
        self.scratch = torch.empty(index.scratch_bytes, dtype=torch.uint8, device=index.device)
        self.topk_ws = Workspace(index.device)

    def search(self, request) -> SearchHits:
        r = request if isinstance(request, _Request) else \
            pack_request(request, self.ix.mz_power, self.ix.intensity_power)
        if r.n == 0:
            z = np.zeros((0, self.k))
            return SearchHits(z.astype(np.float32), z.astype(np.int32), z.astype(np.int32))
        dev = self.ix.device
        refs = _DeviceRequest(r, dev)
        score = torch.empty((r.n, self.ix.n), dtype=torch.float32, device=dev)
        matches = torch.empty((r.n, self.ix.n), dtype=torch.int32, device=dev)
        overflow = torch.empty((r.n, self.ix.n), dtype=torch.uint8, device=dev)
        self.ix.search(refs.struct(), r.n, score, matches, overflow, self.scratch)
        ts, ti, tm = topk_rows(score, self.k, matches, workspace=self.topk_ws)
        return SearchHits(ts.cpu().numpy(), ti.cpu().numpy(), tm.cpu().numpy())


class _DeviceRequest:
    """One packed request uploaded on its own (the plain path)."""

    def __init__(self, r: _Request, dev):
        def up(a, dt):
            a = np.ascontiguousarray(a, dt)
            return torch.from_numpy(a if a.size else np.zeros(1, dt)).to(dev)

        self.r = r
        self.off = up(r.offsets, np.int64)
        self.mz = up(r.mz32, np.float32)
        self.mzx = None if r.exact else up(r.mz64, np.float64)
        self.mzp = None if r.mz_pow is None else up(r.mz_pow, np.float64)
        self.inp = up(r.int_pow, np.float64)
        self.norm = up(r.norm, np.float64)

    def struct(self):
        r = self.r

        def ptr(t):
            return None if t is None else t.data_ptr()

        return _native.CsSpectra(r.n, r.peaks, ptr(self.off), ptr(self.mz), ptr(self.mzx),
                                 ptr(self.mzp), ptr(self.inp), ptr(self.norm), r.max_peaks,
                                 _f32_up(r.max_abs_mz))


def run_load(kind: str, index: LibraryIndex, requests: Sequence[_Request], arrivals: np.ndarray,
             k: int = 10, **server_kw) -> dict:
    """Replay a request stream (arrival times in seconds from the start) on
    ``kind`` = "served" or "plain"; latency = completion - scheduled arrival.
    The server (slots, streams, pinned staging) or the plain searcher is set
    up before the stream's clock starts: setting up is not request latency,
    and allocating a few GB of slots can take longer than the first arrivals'
    spacing."""
    # objects that exist before the stream starts are moved out of the
    # collector's generations, so a full collection during the stream scans
    # only what the stream itself allocates (a long-lived server does the same
    # after start-up)
    gc.collect()
    gc.freeze()
    try:
        return _run_load(kind, index, requests, arrivals, k, **server_kw)
    finally:
        gc.unfreeze()


def _run_load(kind, index, requests, arrivals, k, **server_kw) -> dict:
    done: List[_Request] = []
    t_setup = time.perf_counter()
    if kind == "served":
        srv = SearchServer(index, k=k, **server_kw).start()
        torch.cuda.synchronize(index.device)
        setup_s = time.perf_counter() - t_setup
        t0 = time.perf_counter() + 0.005
        try:
            futs = []
            for r, a in zip(requests, arrivals):
                now = time.perf_counter()
                if t0 + a > now:
                    time.sleep(t0 + a - now)
                futs.append(srv.submit(r, t_arrival=t0 + a))
            for f in futs:
                f.result()
        finally:
            srv.stop()
        done = list(requests)
        extra = {"batches": srv.batches, "mean_batch_spectra":
                 float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0,
                 "slots": len(srv.slots), "slot_spectra": srv.slot_spectra}
    elif kind == "plain":
        plain = PlainSearch(index, k)
        q: "queue.Queue[Optional[_Request]]" = queue.Queue()
        err: List[BaseException] = []
        torch.cuda.synchronize(index.device)

        def worker():
            while True:
                r = q.get()
                if r is None:
                    break
                try:
                    r.future.set_result(plain.search(r))
                except BaseException as e:
                    err.append(e)
                    r.future.set_exception(e)
                r.t_done = time.perf_counter()

        th = threading.Thread(target=worker, daemon=True)
        th.start()
        setup_s = time.perf_counter() - t_setup
        t0 = time.perf_counter() + 0.005
        for r, a in zip(requests, arrivals):
            now = time.perf_counter()
            if t0 + a > now:
                time.sleep(t0 + a - now)
            r.future = Future()
            r.t_arrival = t0 + a
            q.put(r)
        q.put(None)
        th.join()
        if err:
            raise err[0]
        done = list(requests)
        extra = {}
    else:
        raise ValueError(kind)
    lat = np.array([r.t_done - r.t_arrival for r in done]) * 1e3
    span = max(r.t_done for r in done) - t0 if done else 0.0
    nspec = sum(r.n for r in done)
    return {"mode": kind, "requests": len(done), "query_spectra": nspec,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
            "p999_ms": float(np.percentile(lat, 99.9)) if len(lat) else 0.0,
            "max_ms": float(lat.max()) if len(lat) else 0.0,
            "mean_ms": float(lat.mean()) if len(lat) else 0.0,
            "setup_ms": setup_s * 1e3,
            "requests_per_s": len(done) / span if span > 0 else 0.0,
            "query_spectra_per_s": nspec / span if span > 0 else 0.0, **extra}
