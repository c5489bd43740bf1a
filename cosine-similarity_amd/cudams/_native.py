"""ctypes binding of libcsamd.so (the gfx950 HIP library, C ABI in
include/cosine_similarity_amd.h).

There is no fallback: if the library is missing or the machine has no ROCm
GPU, every compute call raises.  Build it with
``python cosine-similarity_amd/build.py``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsamd.so")
CPU_LIB_PATH = os.path.join(_HERE, "libcsamd_cpu.so")

CS_OK = 0
CS_ERR_INVALID = 1
CS_ERR_WORKSPACE = 2
CS_ERR_HIP = 3
CS_ERR_UNSUPPORTED = 4

CS_JACCARD, CS_DICE, CS_COSINE = 0, 1, 2

# Every symbol include/cosine_similarity_amd.h declares.
EXPORTED_SYMBOLS = (
    "cs_last_error",
    "cs_version",
    "cs_cosine_greedy_workspace_size",
    "cs_cosine_greedy",
    "cs_cosine_greedy_indexed_sizes",
    "cs_cosine_greedy_indexed",
    "cs_cosine_batch_create",
    "cs_cosine_batch_upload",
    "cs_cosine_batch_run",
    "cs_cosine_batch_destroy",
    "cs_fingerprint_workspace_size",
    "cs_fingerprint_similarity",
    "cs_topk_rows_workspace_size",
    "cs_topk_rows",
    "cs_threshold_compact",
    "cs_validate_spectra",
    "cs_spectra_norms",
    "cs_cosine_greedy_noalloc",
    "cs_pack_records",
)

NOALLOC_WORKSPACE_BYTES = 64 << 20  # CS_NOALLOC_WORKSPACE_BYTES

CS_FAULT_OFFSETS, CS_FAULT_LENGTH, CS_FAULT_MZ_RANGE, CS_FAULT_MZ_ORDER = 1, 2, 4, 8
FAULT_NAMES = {CS_FAULT_OFFSETS: "offsets decrease, are negative or pass n_peaks",
               CS_FAULT_LENGTH: "a spectrum has more peaks than max_peaks",
               CS_FAULT_MZ_RANGE: "an m/z is not finite or exceeds max_abs_mz",
               CS_FAULT_MZ_ORDER: "m/z not ascending within a spectrum"}


class CsSpectra(ctypes.Structure):
    """struct cs_spectra."""

    _fields_ = [
        ("n_spectra", ctypes.c_int64),
        ("n_peaks", ctypes.c_int64),
        ("offsets", ctypes.c_void_p),
        ("mz", ctypes.c_void_p),
        ("mz_exact", ctypes.c_void_p),
        ("mz_pow", ctypes.c_void_p),
        ("int_pow", ctypes.c_void_p),
        ("norm", ctypes.c_void_p),
        ("max_peaks", ctypes.c_int32),
        ("max_abs_mz", ctypes.c_float),
    ]


class CsCosineParams(ctypes.Structure):
    """struct cs_cosine_params."""

    _fields_ = [
        ("tolerance", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("match_limit", ctypes.c_int64),
        ("light_candidates", ctypes.c_int32),
        ("round_capacity", ctypes.c_int32),
        ("refs_per_block", ctypes.c_int32),
        ("stages", ctypes.c_int32),
    ]


CS_STAGE_PREPARE, CS_STAGE_LIGHT, CS_STAGE_HEAVY = 1, 2, 4


class CsCosineJob(ctypes.Structure):
    """struct cs_cosine_job."""

    _fields_ = [
        ("refs", CsSpectra),
        ("queries", CsSpectra),
        ("tolerance", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("match_limit", ctypes.c_int64),
        ("score", ctypes.c_void_p),
        ("matches", ctypes.c_void_p),
        ("overflow", ctypes.c_void_p),
        ("ld", ctypes.c_int64),
    ]


class CsBatchParams(ctypes.Structure):
    """struct cs_batch_params."""

    _fields_ = [
        ("light_candidates", ctypes.c_int32),
        ("round_capacity", ctypes.c_int32),
        ("refs_per_block", ctypes.c_int32),
        ("grid", ctypes.c_int32),
    ]


class NativeError(RuntimeError):
    """A cs_* call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"libcsamd status {status}: {message}")
        self.status = status


_lib = None


def load(path: str = LIB_PATH):
    """Load libcsamd.so (once) and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python cosine-similarity_amd/build.py` "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    # torch's wheel carries its own HIP runtime.  Loaded first, it is the one
    # the library binds to (same soname); loaded after the library, it would
    # be a second runtime in the process, and whichever opens the GPU second
    # finds no device.
    import torch  # noqa: F401
    L = ctypes.CDLL(path)
    p, i64, i32, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
    L.cs_last_error.restype = ctypes.c_char_p
    L.cs_last_error.argtypes = []
    L.cs_version.restype = ctypes.c_int
    L.cs_version.argtypes = []
    L.cs_cosine_greedy_workspace_size.restype = i32
    L.cs_cosine_greedy_workspace_size.argtypes = [ctypes.POINTER(CsSpectra),
                                                  ctypes.POINTER(CsSpectra),
                                                  ctypes.POINTER(sz)]
    L.cs_cosine_greedy.restype = i32
    L.cs_cosine_greedy.argtypes = [ctypes.POINTER(CsSpectra), ctypes.POINTER(CsSpectra),
                                   ctypes.POINTER(CsCosineParams), p, p, p, i64, p, sz, p]
    L.cs_fingerprint_workspace_size.restype = i32
    L.cs_fingerprint_workspace_size.argtypes = [i64, i64, i64, ctypes.POINTER(sz)]
    L.cs_fingerprint_similarity.restype = i32
    L.cs_fingerprint_similarity.argtypes = [i32, p, i64, p, i64, i64, p, p, ctypes.c_float, p,
                                            p, i64, p, sz, p]
    L.cs_cosine_greedy_indexed_sizes.restype = i32
    L.cs_cosine_greedy_indexed_sizes.argtypes = [ctypes.POINTER(CsSpectra), ctypes.POINTER(sz),
                                                 ctypes.POINTER(sz)]
    L.cs_cosine_greedy_indexed.restype = i32
    L.cs_cosine_greedy_indexed.argtypes = [ctypes.POINTER(CsSpectra), ctypes.POINTER(CsSpectra),
                                           ctypes.POINTER(CsCosineParams), p, p, p, i64, p, sz, p,
                                           sz, p]
    L.cs_topk_rows_workspace_size.restype = i32
    L.cs_topk_rows_workspace_size.argtypes = [i64, i64, i32, ctypes.POINTER(sz)]
    L.cs_topk_rows.restype = i32
    L.cs_topk_rows.argtypes = [p, p, i64, i64, i64, i32, p, p, p, p, sz, p]
    L.cs_cosine_batch_create.restype = i32
    L.cs_cosine_batch_create.argtypes = [ctypes.POINTER(CsCosineJob), ctypes.c_int32,
                                         ctypes.POINTER(CsBatchParams),
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]
    L.cs_cosine_batch_upload.restype = i32
    L.cs_cosine_batch_upload.argtypes = [p, p, sz, p]
    L.cs_cosine_batch_run.restype = i32
    L.cs_cosine_batch_run.argtypes = [p, p]
    L.cs_cosine_batch_destroy.restype = None
    L.cs_cosine_batch_destroy.argtypes = [p]
    L.cs_threshold_compact.restype = i32
    L.cs_threshold_compact.argtypes = [p, p, p, i64, i64, i64, ctypes.c_float, i64, p, p, i64,
                                       p, p, p, p, p, p]
    L.cs_cosine_greedy_noalloc.restype = i32
    L.cs_cosine_greedy_noalloc.argtypes = [ctypes.POINTER(CsSpectra), ctypes.POINTER(CsSpectra),
                                           ctypes.POINTER(CsCosineParams), p, p, p, i64, p]
    L.cs_pack_records.restype = i32
    L.cs_pack_records.argtypes = [p, p, p, i64, i64, i64, p, p]
    L.cs_validate_spectra.restype = i32
    L.cs_validate_spectra.argtypes = [ctypes.POINTER(CsSpectra), p, p]
    L.cs_spectra_norms.restype = i32
    L.cs_spectra_norms.argtypes = [ctypes.POINTER(CsSpectra), p, p]
    _lib = L
    return L


# Every symbol of the host library (libcsamd_cpu.so).
CPU_EXPORTED_SYMBOLS = ("cs_cpu_last_error", "cs_cpu_cosine_greedy")

_cpu_lib = None


def load_cpu(path: str = CPU_LIB_PATH):
    """Load libcsamd_cpu.so (the host engine of CPUParallelCosineGreedy)."""
    global _cpu_lib
    if _cpu_lib is not None:
        return _cpu_lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python cosine-similarity_amd/build.py`.")
    L = ctypes.CDLL(path)
    p = ctypes.c_void_p
    L.cs_cpu_last_error.restype = ctypes.c_char_p
    L.cs_cpu_last_error.argtypes = []
    L.cs_cpu_cosine_greedy.restype = ctypes.c_int
    L.cs_cpu_cosine_greedy.argtypes = [ctypes.POINTER(CsSpectra), ctypes.POINTER(CsSpectra),
                                       ctypes.c_double, ctypes.c_double, p, p, ctypes.c_int64,
                                       ctypes.c_int]
    _cpu_lib = L
    return L


def check_cpu(status: int) -> None:
    if status != CS_OK:
        msg = load_cpu().cs_cpu_last_error().decode(errors="replace")
        if status == CS_ERR_INVALID:
            raise ValueError(f"libcsamd_cpu: {msg}")
        raise NativeError(status, msg)


def check(status: int) -> None:
    if status != CS_OK:
        msg = load().cs_last_error().decode(errors="replace")
        if status == CS_ERR_INVALID:
            raise ValueError(f"libcsamd: {msg}")
        raise NativeError(status, msg)
