"""Device-resident entry points: thin wrappers over the C ABI.

Inputs and outputs are torch tensors already on the GPU; calls are
asynchronous on torch's current stream.  The plug-in classes, the batched-job
runner, the server and bench.py all go through here.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _native
from .spectra import DeviceSpectra


class Workspace:
    """A grow-only device byte buffer reused across calls on one stream."""

    def __init__(self, device="cuda"):
        self.device = device
        self.buf: Optional[torch.Tensor] = None

    def get(self, nbytes: int, keep: bool = False) -> torch.Tensor:
        """The buffer, grown to ``nbytes`` if needed.  ``keep=True`` is for calls
        that read state an earlier call left in the buffer: growing would lose
        it, so it raises instead."""
        nbytes = max(int(nbytes), 256)
        if self.buf is None or self.buf.numel() < nbytes:
            if keep:
                raise ValueError("workspace too small for a call that reuses its prepared "
                                 "state; run the PREPARE stage on it first")
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.buf


def _stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU is visible; the HIP library has no CPU path")


def cosine_greedy_workspace_bytes(refs: DeviceSpectra, queries: DeviceSpectra) -> int:
    L = _native.load()
    n = ctypes.c_size_t(0)
    rs, qs = refs.struct(), queries.struct()
    _native.check(L.cs_cosine_greedy_workspace_size(ctypes.byref(rs), ctypes.byref(qs),
                                                    ctypes.byref(n)))
    return int(n.value)


def cosine_greedy(refs: DeviceSpectra, queries: DeviceSpectra, tolerance: float, shift: float,
                  match_limit: int, score: torch.Tensor, matches: torch.Tensor,
                  overflow: torch.Tensor, workspace: Optional[Workspace] = None,
                  light_candidates: int = 0, round_capacity: int = 0, stream=None,
                  stages: int = 0, refs_per_block: int = 0) -> None:
    """All pairs into row-major [R, ld] outputs (float32, int32, uint8).
    ``stages`` (a mask of _native.CS_STAGE_*) runs part of the pipeline."""
    require_gpu()
    L = _native.load()
    R, Q = refs.n_spectra, queries.n_spectra
    for t, dt in ((score, torch.float32), (matches, torch.int32), (overflow, torch.uint8)):
        if t.dtype != dt or not t.is_cuda or t.dim() != 2 or t.shape[0] < R or t.shape[1] < Q:
            raise ValueError(f"output must be a CUDA {dt} tensor of shape >= ({R}, {Q})")
        if t.stride(1) != 1 or t.stride(0) != score.stride(0):
            raise ValueError("outputs must be row-major with one common row stride")
    ws = workspace or Workspace(score.device)
    nbytes = cosine_greedy_workspace_bytes(refs, queries)
    reuse = stages != 0 and not (stages & _native.CS_STAGE_PREPARE)
    if reuse and workspace is None:
        raise ValueError("stages without PREPARE need the workspace a PREPARE call filled")
    buf = ws.get(nbytes, keep=reuse)
    rs, qs = refs.struct(), queries.struct()
    p = _native.CsCosineParams(float(tolerance), float(shift), int(match_limit),
                               int(light_candidates), int(round_capacity), int(refs_per_block),
                               int(stages))
    _native.check(L.cs_cosine_greedy(ctypes.byref(rs), ctypes.byref(qs), ctypes.byref(p),
                                     score.data_ptr(), matches.data_ptr(), overflow.data_ptr(),
                                     score.stride(0), buf.data_ptr(), buf.numel(),
                                     _stream_ptr(stream)))


def cosine_greedy_noalloc(refs, queries, tolerance: float, shift: float, match_limit: int,
                          score: torch.Tensor, matches: torch.Tensor, overflow: torch.Tensor,
                          light_candidates: int = 0, round_capacity: int = 0, stream=None,
                          stages: int = 0) -> None:
    """cs_cosine_greedy_noalloc: as :func:`cosine_greedy`, with the library's
    own fixed device workspace (nothing is allocated; overlapping calls on one
    GPU race)."""
    require_gpu()
    L = _native.load()
    R, Q = refs.n_spectra, queries.n_spectra
    for t, dt in ((score, torch.float32), (matches, torch.int32), (overflow, torch.uint8)):
        if t.dtype != dt or not t.is_cuda or t.dim() != 2 or t.shape[0] < R or t.shape[1] < Q:
            raise ValueError(f"output must be a CUDA {dt} tensor of shape >= ({R}, {Q})")
        if t.stride(1) != 1 or t.stride(0) != score.stride(0):
            raise ValueError("outputs must be row-major with one common row stride")
    rs, qs = refs.struct(), queries.struct()
    p = _native.CsCosineParams(float(tolerance), float(shift), int(match_limit),
                               int(light_candidates), int(round_capacity), 0, int(stages))
    _native.check(L.cs_cosine_greedy_noalloc(ctypes.byref(rs), ctypes.byref(qs), ctypes.byref(p),
                                             score.data_ptr(), matches.data_ptr(),
                                             overflow.data_ptr(), score.stride(0),
                                             _stream_ptr(stream)))


def pack_records(score: torch.Tensor, matches: torch.Tensor, overflow: torch.Tensor,
                 records: torch.Tensor, rows: Optional[int] = None, cols: Optional[int] = None,
                 stream=None) -> None:
    """Interleave the three planes into packed 9-byte records (score f32,
    matches i32, overflow u8) in the uint8 device tensor ``records``."""
    require_gpu()
    L = _native.load()
    rows = score.shape[0] if rows is None else rows
    cols = score.shape[1] if cols is None else cols
    if records.dtype != torch.uint8 or not records.is_contiguous() or records.numel() < 9 * rows * cols:
        raise ValueError("records must be a contiguous uint8 tensor of 9 * rows * cols bytes")
    if matches.stride(0) != score.stride(0) or overflow.stride(0) != score.stride(0):
        raise ValueError("the three planes must share one row stride")
    _native.check(L.cs_pack_records(score.data_ptr(), matches.data_ptr(), overflow.data_ptr(),
                                    rows, cols, score.stride(0), records.data_ptr(),
                                    _stream_ptr(stream)))


def validate_spectra(spectra, stream=None) -> int:
    """CS_FAULT_* bits of a device-resident spectrum set (0 = sound), from the
    device validator cs_validate_spectra; synchronises ``stream``."""
    require_gpu()
    L = _native.load()
    s = spectra.struct()
    dev = spectra.offsets.device if hasattr(spectra, "offsets") else spectra.parent.offsets.device
    flag = torch.empty(1, dtype=torch.int32, device=dev)
    _native.check(L.cs_validate_spectra(ctypes.byref(s), flag.data_ptr(), _stream_ptr(stream)))
    return int(flag.item()) & 0xffffffff


def check_spectra_on_device(spectra, stream=None) -> None:
    """Raises ValueError naming every fault the device validator finds."""
    f = validate_spectra(spectra, stream)
    if f:
        raise ValueError("corrupt spectrum set: " + "; ".join(
            msg for bit, msg in _native.FAULT_NAMES.items() if f & bit))


MEASURES = {"jaccard": _native.CS_JACCARD, "dice": _native.CS_DICE, "cosine": _native.CS_COSINE}


def fingerprint_similarity(measure: str, ref_words: torch.Tensor, query_words: torch.Tensor,
                           out: torch.Tensor, ref_valid: Optional[torch.Tensor] = None,
                           query_valid: Optional[torch.Tensor] = None,
                           empty_score: float = float("nan"),
                           intersections: Optional[torch.Tensor] = None,
                           workspace: Optional[Workspace] = None, stream=None) -> None:
    """out[r, q] = measure between bit-packed fingerprints (uint32 words as int32)."""
    require_gpu()
    L = _native.load()
    if measure not in MEASURES:
        raise ValueError(f"unknown similarity measure {measure!r}")
    R, W = ref_words.shape
    Q, W2 = query_words.shape
    if W != W2:
        raise ValueError("fingerprints differ in length")
    if not (ref_words.is_contiguous() and query_words.is_contiguous()):
        raise ValueError("fingerprint word arrays must be contiguous")
    if out.dtype != torch.float32 or out.shape[0] < R or out.shape[1] < Q or out.stride(1) != 1:
        raise ValueError("out must be a row-major float32 tensor of shape >= (R, Q)")
    ws = workspace or Workspace(out.device)
    n = ctypes.c_size_t(0)
    _native.check(L.cs_fingerprint_workspace_size(R, Q, W, ctypes.byref(n)))
    buf = ws.get(n.value)

    def ptr(t):
        return None if t is None else t.data_ptr()

    if intersections is not None and (intersections.dtype != torch.int32 or
                                      intersections.stride(0) != out.stride(0)):
        raise ValueError("intersections must be int32 with the row stride of out")
    _native.check(L.cs_fingerprint_similarity(
        MEASURES[measure], ref_words.data_ptr(), R,
        query_words.data_ptr(), Q, W, ptr(ref_valid), ptr(query_valid),
        float(empty_score), out.data_ptr(), ptr(intersections), out.stride(0), buf.data_ptr(),
        buf.numel(), _stream_ptr(stream)))


def topk_rows(score: torch.Tensor, k: int, matches: Optional[torch.Tensor] = None,
              rows: Optional[int] = None, cols: Optional[int] = None, stream=None,
              workspace: Optional[Workspace] = None, out=None):
    """Per row, the k best (score desc, column asc): (scores [rows, k] f32,
    columns [rows, k] i32, matches [rows, k] i32 or None).  ``out`` may hold
    three preallocated tensors to write into (the third None without matches)."""
    require_gpu()
    L = _native.load()
    rows = score.shape[0] if rows is None else rows
    cols = score.shape[1] if cols is None else cols
    if score.dtype != torch.float32 or score.stride(1) != 1:
        raise ValueError("score must be a row-major float32 tensor")
    if matches is not None and matches.stride(0) != score.stride(0):
        raise ValueError("matches must share the row stride of score")
    if out is None:
        ts = torch.empty((rows, k), dtype=torch.float32, device=score.device)
        ti = torch.empty((rows, k), dtype=torch.int32, device=score.device)
        tm = None if matches is None else torch.empty((rows, k), dtype=torch.int32,
                                                      device=score.device)
    else:
        ts, ti, tm = out
        for t, dt in ((ts, torch.float32), (ti, torch.int32), (tm, torch.int32)):
            if t is not None and (t.dtype != dt or not t.is_contiguous() or t.numel() < rows * k):
                raise ValueError("top-k outputs must be contiguous with rows * k elements")
    n = ctypes.c_size_t(0)
    _native.check(L.cs_topk_rows_workspace_size(rows, cols, int(k), ctypes.byref(n)))
    ws = workspace or Workspace(score.device)
    buf = ws.get(n.value)
    _native.check(L.cs_topk_rows(score.data_ptr(), None if matches is None else matches.data_ptr(),
                                 rows, cols, score.stride(0), int(k), ts.data_ptr(), ti.data_ptr(),
                                 None if tm is None else tm.data_ptr(), buf.data_ptr(), buf.numel(),
                                 _stream_ptr(stream)))
    return ts, ti, tm


def cosine_greedy_indexed_bytes(queries: DeviceSpectra):
    """(index bytes, scratch bytes) of cs_cosine_greedy_indexed for these queries."""
    L = _native.load()
    a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
    qs = queries.struct()
    _native.check(L.cs_cosine_greedy_indexed_sizes(ctypes.byref(qs), ctypes.byref(a),
                                                   ctypes.byref(b)))
    return int(a.value), int(b.value)


def cosine_greedy_indexed(refs: DeviceSpectra, queries: DeviceSpectra, tolerance: float,
                          shift: float, match_limit: int, score: Optional[torch.Tensor],
                          matches: Optional[torch.Tensor], overflow: Optional[torch.Tensor],
                          index: torch.Tensor, scratch: Optional[torch.Tensor], stages: int = 0,
                          light_candidates: int = 0, round_capacity: int = 0,
                          stream=None, refs_per_block: int = 0) -> None:
    """cs_cosine_greedy_indexed: the query index lives in ``index`` (filled by a
    PREPARE call, then shared read-only), the call's own scratch in
    ``scratch``.  Outputs may be None for a PREPARE-only call."""
    require_gpu()
    L = _native.load()
    R, Q = refs.n_spectra, queries.n_spectra
    computes = stages == 0 or (stages & (_native.CS_STAGE_LIGHT | _native.CS_STAGE_HEAVY))
    ld = 0
    if computes:
        for t, dt in ((score, torch.float32), (matches, torch.int32), (overflow, torch.uint8)):
            if (t is None or t.dtype != dt or not t.is_cuda or t.dim() != 2 or t.shape[0] < R
                    or t.shape[1] < Q):
                raise ValueError(f"output must be a CUDA {dt} tensor of shape >= ({R}, {Q})")
            if t.stride(1) != 1 or t.stride(0) != score.stride(0):
                raise ValueError("outputs must be row-major with one common row stride")
        ld = score.stride(0)
    rs, qs = refs.struct(), queries.struct()
    p = _native.CsCosineParams(float(tolerance), float(shift), int(match_limit),
                               int(light_candidates), int(round_capacity), int(refs_per_block),
                               int(stages))

    def ptr(t):
        return None if t is None else t.data_ptr()

    _native.check(L.cs_cosine_greedy_indexed(
        ctypes.byref(rs), ctypes.byref(qs), ctypes.byref(p), ptr(score), ptr(matches),
        ptr(overflow), ld, index.data_ptr(), index.numel(), ptr(scratch),
        0 if scratch is None else scratch.numel(), _stream_ptr(stream)))


def threshold_compact(score: torch.Tensor, matches: torch.Tensor, overflow: torch.Tensor,
                      threshold: float, row0: int = 0, rows: Optional[int] = None,
                      cols: Optional[int] = None, stream=None):
    """COO of entries with score >= threshold in row-major order, copied to the
    host: (row i64, col i64, score f32, matches i32, overflow u8)."""
    require_gpu()
    L = _native.load()
    rows = score.shape[0] if rows is None else rows
    cols = score.shape[1] if cols is None else cols
    dev = score.device
    ld = score.stride(0)
    offs = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    counts = torch.empty(max(rows, 1), dtype=torch.int32, device=dev)
    sp = _stream_ptr(stream)
    _native.check(L.cs_threshold_compact(score.data_ptr(), matches.data_ptr(), overflow.data_ptr(),
                                         rows, cols, ld, float(threshold), int(row0),
                                         offs.data_ptr(), counts.data_ptr(), 0, None, None, None,
                                         None, None, sp))
    total = int(offs[rows].item()) if rows else 0
    n = max(total, 1)
    r = torch.empty(n, dtype=torch.int64, device=dev)
    c = torch.empty(n, dtype=torch.int64, device=dev)
    s = torch.empty(n, dtype=torch.float32, device=dev)
    m = torch.empty(n, dtype=torch.int32, device=dev)
    o = torch.empty(n, dtype=torch.uint8, device=dev)
    if total:
        _native.check(L.cs_threshold_compact(score.data_ptr(), matches.data_ptr(),
                                             overflow.data_ptr(), rows, cols, ld, float(threshold),
                                             int(row0), offs.data_ptr(), counts.data_ptr(), total,
                                             r.data_ptr(), c.data_ptr(), s.data_ptr(),
                                             m.data_ptr(), o.data_ptr(), sp))
    return tuple(x[:total].cpu().numpy() for x in (r, c, s, m, o))
