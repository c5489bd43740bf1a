"""CPU oracle for the spectral-similarity path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or as the timed
CPU baseline).  The product path in ``cosine-similarity_amd/`` never imports
it and fails loudly when its HIP library is missing.

The arithmetic lives in ``cs_oracle.c`` (float64 restatement of matchms
CosineGreedy / FingerprintSimilarity as the reference calls them, see
cudams/similarity/spectrum_similarity_functions.py:306-317 and
cudams/similarity/vector_similarity_functions.py:42-46, :87-90, :116-118).
This module only loads it with ctypes and packs numpy inputs.

``greedy_pair_py`` is a second, literal pure-Python restatement used for tiny
brute-force cross-checks: it derives the greedy order from an explicit
(value desc, i desc, j desc) key instead of the stable-sort-then-reverse the C
code follows, so agreement pins that the two orderings are the same.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcs_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build() -> str:
    """Compile cs_oracle.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.cso_cosine_greedy_pair.argtypes = [_p, _p, _i64, _p, _p, _i64, _d, _d, _d, _d, _p, _p, _p]
        L.cso_cosine_greedy_pair_assign.argtypes = [_p, _p, _i64, _p, _p, _i64, _d, _d, _d, _d,
                                                    _p, _p, _i64, _p, _p]
        L.cso_cosine_greedy_matrix.argtypes = [_p, _p, _p, _i64, _p, _p, _p, _i64, _d, _d, _d, _d,
                                               _i64, _p, _p, _p, _p, ctypes.c_int]
        L.cso_near_tie_pairs.argtypes = [_p, _p, _p, _i64, _p, _p, _p, _i64, _d, _d, _d, _d, _p,
                                         ctypes.c_int]
        L.cso_near_tie_pairs.restype = _i64
        L.cso_fingerprint_matrix.argtypes = [_p, _i64, _p, _i64, _i64, ctypes.c_int, _p, _p,
                                             ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def cosine_greedy_pair(mz1, int1, mz2, int2, tolerance=0.1, shift=0.0, mz_power=0.0,
                       intensity_power=1.0) -> Tuple[float, int, int]:
    """(score, matches, n_candidates) of one (reference, query) pair."""
    a, b, c, d = _f64(mz1), _f64(int1), _f64(mz2), _f64(int2)
    s = np.zeros(1, np.float64)
    m = np.zeros(1, np.int32)
    n = np.zeros(1, np.int64)
    lib().cso_cosine_greedy_pair(_ptr(a), _ptr(b), len(a), _ptr(c), _ptr(d), len(c), tolerance,
                                 shift, mz_power, intensity_power, _ptr(s), _ptr(m), _ptr(n))
    return float(s[0]), int(m[0]), int(n[0])


def cosine_greedy_assignment(mz1, int1, mz2, int2, tolerance=0.1, shift=0.0, mz_power=0.0,
                             intensity_power=1.0):
    """(score, [(i, j) accepted in greedy order])."""
    a, b, c, d = _f64(mz1), _f64(int1), _f64(mz2), _f64(int2)
    cap = max(1, min(len(a), len(c)))
    ai = np.zeros(cap, np.int32)
    aj = np.zeros(cap, np.int32)
    s = np.zeros(1, np.float64)
    m = np.zeros(1, np.int32)
    lib().cso_cosine_greedy_pair_assign(_ptr(a), _ptr(b), len(a), _ptr(c), _ptr(d), len(c),
                                        tolerance, shift, mz_power, intensity_power, _ptr(ai),
                                        _ptr(aj), cap, _ptr(s), _ptr(m))
    k = int(m[0])
    return float(s[0]), list(zip(ai[:k].tolist(), aj[:k].tolist()))


def pack(spectra: Sequence[Tuple[np.ndarray, np.ndarray]]):
    """List of (mz, intensity) -> float64 CSR (mz, intensity, offsets)."""
    lens = np.array([len(s[0]) for s in spectra], dtype=np.int64)
    off = np.zeros(len(spectra) + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    if len(spectra) and off[-1] > 0:
        mz = np.concatenate([_f64(s[0]) for s in spectra])
        it = np.concatenate([_f64(s[1]) for s in spectra])
    else:
        mz = np.zeros(1, np.float64)
        it = np.zeros(1, np.float64)
    return mz, it, off


def cosine_greedy_matrix(refs, queries, tolerance=0.1, shift=0.0, mz_power=0.0,
                         intensity_power=1.0, match_limit=2048, nthreads=0):
    """All-pairs oracle.  refs/queries: lists of (mz, intensity) or CSR tuples.

    Returns dict of score (f64), matches (i32), overflow (u8), n_candidates (i64),
    each [R, Q].
    """
    rmz, rin, roff = refs if isinstance(refs, tuple) and len(refs) == 3 else pack(refs)
    qmz, qin, qoff = queries if isinstance(queries, tuple) and len(queries) == 3 else pack(queries)
    rmz, rin, qmz, qin = _f64(rmz), _f64(rin), _f64(qmz), _f64(qin)
    roff = np.ascontiguousarray(roff, np.int64)
    qoff = np.ascontiguousarray(qoff, np.int64)
    R, Q = len(roff) - 1, len(qoff) - 1
    score = np.zeros((R, Q), np.float64)
    matches = np.zeros((R, Q), np.int32)
    overflow = np.zeros((R, Q), np.uint8)
    ncand = np.zeros((R, Q), np.int64)
    lib().cso_cosine_greedy_matrix(_ptr(rmz), _ptr(rin), _ptr(roff), R, _ptr(qmz), _ptr(qin),
                                   _ptr(qoff), Q, tolerance, shift, mz_power, intensity_power,
                                   int(match_limit), _ptr(score), _ptr(matches), _ptr(overflow),
                                   _ptr(ncand), int(nthreads))
    return {"score": score, "matches": matches, "overflow": overflow, "n_candidates": ncand}


def near_tie_pairs(refs, queries, tolerance=0.1, shift=0.0, mz_power=0.0, intensity_power=1.0,
                   nthreads=0):
    """uint8 [R, Q]: 1 where a greedy over candidate values rounded to float32
    (the reference GPU kernel's stored products) accepts a different set of
    peak pairs than the float64 order of matchms."""
    rmz, rin, roff = refs if isinstance(refs, tuple) and len(refs) == 3 else pack(refs)
    qmz, qin, qoff = queries if isinstance(queries, tuple) and len(queries) == 3 else pack(queries)
    rmz, rin, qmz, qin = _f64(rmz), _f64(rin), _f64(qmz), _f64(qin)
    roff = np.ascontiguousarray(roff, np.int64)
    qoff = np.ascontiguousarray(qoff, np.int64)
    R, Q = len(roff) - 1, len(qoff) - 1
    flags = np.zeros((R, Q), np.uint8)
    lib().cso_near_tie_pairs(_ptr(rmz), _ptr(rin), _ptr(roff), R, _ptr(qmz), _ptr(qin),
                             _ptr(qoff), Q, tolerance, shift, mz_power, intensity_power,
                             _ptr(flags), int(nthreads))
    return flags


def greedy_accepted_py(mz1, int1, mz2, int2, tolerance=0.1, key=float):
    """Pure-Python accepted set {(i, j)} of the greedy with candidate values
    mapped by ``key`` before ordering (v desc, i desc, j desc)."""
    cands = []
    for i in range(len(mz1)):
        for j in range(len(mz2)):
            if mz1[i] - tolerance <= mz2[j] <= mz1[i] + tolerance:
                cands.append((key(float(int1[i]) * float(int2[j])), i, j))
    cands.sort(reverse=True)
    u1, u2, acc = set(), set(), set()
    for _, i, j in cands:
        if i not in u1 and j not in u2:
            u1.add(i)
            u2.add(j)
            acc.add((i, j))
    return acc


MEASURES = {"jaccard": 0, "dice": 1, "cosine": 2}


def fingerprint_matrix(refs: np.ndarray, queries: np.ndarray, measure: str, nthreads=0):
    """(scores f64 [R, Q], intersections i64 [R, Q]) for 0/1 fingerprints [rows, nbits]."""
    r = np.ascontiguousarray((np.asarray(refs) != 0).astype(np.uint8))
    q = np.ascontiguousarray((np.asarray(queries) != 0).astype(np.uint8))
    assert r.ndim == 2 and q.ndim == 2 and r.shape[1] == q.shape[1]
    out = np.zeros((r.shape[0], q.shape[0]), np.float64)
    inter = np.zeros((r.shape[0], q.shape[0]), np.int64)
    lib().cso_fingerprint_matrix(_ptr(r), r.shape[0], _ptr(q), q.shape[0], r.shape[1],
                                 MEASURES[measure], _ptr(out), _ptr(inter), int(nthreads))
    return out, inter


def greedy_pair_py(mz1, int1, mz2, int2, tolerance=0.1, shift=0.0, mz_power=0.0,
                   intensity_power=1.0):
    """Pure-Python brute force for tiny inputs: every (i, j) is tested directly
    (no moving lower pointer), and the greedy order comes from an explicit
    (v desc, i desc, j desc) key.  Returns (score, matches, n_candidates)."""
    mz1, int1, mz2, int2 = (list(map(float, x)) for x in (mz1, int1, mz2, int2))
    cands = []
    for i in range(len(mz1)):
        for j in range(len(mz2)):
            x = mz2[j] + shift
            if (mz1[i] - tolerance) <= x <= (mz1[i] + tolerance):
                v = (mz1[i] ** mz_power) * (int1[i] ** intensity_power)
                v = v * (mz2[j] ** mz_power)
                v = v * (int2[j] ** intensity_power)
                cands.append((v, i, j))
    if not cands:
        return 0.0, 0, 0
    cands.sort(key=lambda t: (t[0], t[1], t[2]), reverse=True)
    u1, u2, s, m = set(), set(), 0.0, 0
    for v, i, j in cands:
        if i not in u1 and j not in u2:
            u1.add(i)
            u2.add(j)
            s += v
            m += 1
    n1 = sum(((a ** mz_power) * (b ** intensity_power)) ** 2 for a, b in zip(mz1, int1)) ** 0.5
    n2 = sum(((a ** mz_power) * (b ** intensity_power)) ** 2 for a, b in zip(mz2, int2)) ** 0.5
    return s / (n1 * n2), m, len(cands)
