"""Spectra -> ragged CSR, host and device side.

Replaces the padded ``[2, batch, n_max_peaks]`` float32 packing of the
reference (cudams/similarity/CudaCosineGreedy.py:63-87, cudams/utils.py:60-91):
no padding, one copy per spectrum set, and the float64 values matchms works
with are kept next to their float32 rounding.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

MAX_PEAKS = 65535  # peak indices travel as 16-bit halves of one 32-bit key


def spectrum_peaks(s) -> Tuple[np.ndarray, np.ndarray]:
    """(mz, intensities) as float64 arrays of one spectrum-like object.

    Accepts matchms ``Spectrum`` (``.peaks.mz`` / ``.peaks.intensities``), this
    package's :class:`cudams.spectrum.Spectrum`, a ``(mz, intensities)`` pair,
    or ``None`` (an empty spectrum).
    """
    if s is None:
        return np.zeros(0, np.float64), np.zeros(0, np.float64)
    peaks = getattr(s, "peaks", None)
    if peaks is not None:
        return (np.asarray(peaks.mz, dtype=np.float64),
                np.asarray(peaks.intensities, dtype=np.float64))
    if isinstance(s, (tuple, list)) and len(s) == 2:
        return np.asarray(s[0], dtype=np.float64), np.asarray(s[1], dtype=np.float64)
    raise TypeError(f"cannot read peaks from {type(s).__name__}")


@dataclass
class PackedSpectra:
    """Host CSR: spectrum s owns peaks [offsets[s], offsets[s+1])."""

    mz: np.ndarray           # float64 [P]
    intensities: np.ndarray  # float64 [P]
    offsets: np.ndarray      # int64 [n + 1]

    @property
    def n_spectra(self) -> int:
        return len(self.offsets) - 1

    @property
    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    @classmethod
    def from_spectra(cls, spectra: Sequence, n_max_peaks: Optional[int] = None) -> "PackedSpectra":
        """Pack spectra, keeping the first ``n_max_peaks`` peaks of each (as the
        reference does at CudaCosineGreedy.py:82)."""
        pairs = [spectrum_peaks(s) for s in spectra]
        mzs = [mz for mz, _ in pairs]
        its = [it for _, it in pairs]
        if any(mz.ndim != 1 or it.ndim != 1 or len(mz) != len(it) for mz, it in pairs):
            raise ValueError("m/z and intensity arrays must be 1-D of equal length")
        if n_max_peaks is not None:
            for k in np.flatnonzero(np.fromiter(map(len, mzs), np.int64, len(mzs)) > n_max_peaks):
                mzs[k], its[k] = mzs[k][:n_max_peaks], its[k][:n_max_peaks]
        return cls._concat(mzs, its)  # already float64 1-D arrays

    @classmethod
    def from_arrays(cls, mzs: Iterable[np.ndarray], intensities: Iterable[np.ndarray]) -> "PackedSpectra":
        return cls._concat([np.asarray(m, np.float64) for m in mzs],
                           [np.asarray(i, np.float64) for i in intensities])

    @classmethod
    def _concat(cls, mzs: list, its: list) -> "PackedSpectra":
        lens = np.fromiter(map(len, mzs), dtype=np.int64, count=len(mzs))
        off = np.zeros(len(mzs) + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        mz = np.concatenate(mzs) if len(mzs) else np.zeros(0, np.float64)
        it = np.concatenate(its) if len(its) else np.zeros(0, np.float64)
        p = cls(np.ascontiguousarray(mz), np.ascontiguousarray(it), off)
        p.validate()
        return p

    def validate(self) -> None:
        off = self.offsets
        if off[0] != 0 or np.any(np.diff(off) < 0) or off[-1] != len(self.mz):
            raise ValueError("corrupt CSR offsets")
        if len(self.mz) != len(self.intensities):
            raise ValueError("m/z and intensity arrays differ in length")
        if len(self.mz):
            if not (np.isfinite(self.mz).all() and np.isfinite(self.intensities).all()):
                raise ValueError("m/z and intensities must be finite")
            # a decrease of m/z is allowed only from a spectrum's last peak
            # to the next spectrum's first
            down = np.flatnonzero(self.mz[1:] < self.mz[:-1])
            if down.size:
                start = np.zeros(len(self.mz) + 1, dtype=bool)
                start[off[1:-1]] = True
                if not start[down + 1].all():
                    raise ValueError("peaks must be sorted by m/z within each spectrum")
        if self.n_spectra and int(self.lengths.max()) > MAX_PEAKS:
            raise ValueError(f"a spectrum has more than {MAX_PEAKS} peaks; set n_max_peaks")


def peak_factors(p: PackedSpectra, mz_power: float, intensity_power: float):
    """Per-peak float64 factors as matchms writes them: (mz ** mz_power, or
    None when mz_power == 0, every factor then being exactly 1.0;
    intensities ** intensity_power)."""
    mz_pow = None if mz_power == 0 else p.mz ** mz_power
    # x ** 1.0 is x exactly (IEEE pow); skip the pass for the default power
    int_pow = p.intensities if intensity_power == 1 else p.intensities ** intensity_power
    return mz_pow, np.ascontiguousarray(int_pow, np.float64)


def power_factors(p: PackedSpectra, mz_power: float, intensity_power: float):
    """peak_factors and the per-spectrum norms sqrt(sum((mz**p * I**q) ** 2)),
    summed left to right (matchms score_best_matches; the reference GPU kernel
    computes the same norms at spectrum_similarity_functions.py:82-110).
    cs_spectra_norms computes the same values on the device."""
    mz_pow, int_pow = peak_factors(p, mz_power, intensity_power)
    w = int_pow if mz_pow is None else mz_pow * int_pow
    seg = np.repeat(np.arange(p.n_spectra), p.lengths)
    norm = np.sqrt(np.bincount(seg, weights=w * w, minlength=p.n_spectra)) if p.n_spectra else \
        np.zeros(0, np.float64)
    return mz_pow, int_pow, norm.astype(np.float64)


class DeviceSpectra:
    """A PackedSpectra resident in device memory, prepared for one
    (mz_power, intensity_power), exposing the C struct ``cs_spectra``."""

    def __init__(self, packed: PackedSpectra, mz_power: float = 0.0, intensity_power: float = 1.0,
                 device="cuda"):
        import torch

        self.packed = packed
        self.mz_power, self.intensity_power = float(mz_power), float(intensity_power)
        mz_pow, int_pow = peak_factors(packed, mz_power, intensity_power)
        mz32 = packed.mz.astype(np.float32)
        exact = np.array_equal(mz32.astype(np.float64), packed.mz)

        def up(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            return torch.from_numpy(a).to(device)

        self.offsets = up(packed.offsets, np.int64)
        self.mz = up(mz32, np.float32)
        self.mz_exact = None if exact else up(packed.mz, np.float64)
        self.mz_pow = None if mz_pow is None else up(mz_pow, np.float64)
        self.int_pow = up(int_pow, np.float64)
        self.n_spectra = packed.n_spectra
        self.n_peaks = int(packed.offsets[-1])
        self.max_peaks = int(packed.lengths.max()) if self.n_spectra else 0
        self.max_abs_mz = float(max(-packed.mz.min(), packed.mz.max())) if self.n_peaks else 0.0
        # the norms on the device (cs_spectra_norms: the host's left-to-right
        # sum), on the current stream after the uploads
        from . import _native

        self.norm = torch.empty(max(self.n_spectra, 1), dtype=torch.float64, device=device)
        s = self.struct()
        with torch.cuda.device(self.norm.device):
            _native.check(_native.load().cs_spectra_norms(
                ctypes.byref(s), self.norm.data_ptr(), torch.cuda.current_stream().cuda_stream))

    def slice(self, start: int, stop: int) -> "SpectraSlice":
        """Spectra [start, stop) as a view sharing this object's device memory."""
        if not 0 <= start <= stop <= self.n_spectra:
            raise IndexError("slice outside the spectrum set")
        return SpectraSlice(self, start, stop)

    def struct(self):
        from ._native import CsSpectra

        def ptr(t):
            return None if t is None else t.data_ptr()

        # float32 max |m/z| rounded up so it bounds every float64 value
        m = np.float32(self.max_abs_mz)
        if float(m) < self.max_abs_mz:
            m = np.nextafter(m, np.float32(np.inf))
        return CsSpectra(self.n_spectra, self.n_peaks, ptr(self.offsets), ptr(self.mz),
                         ptr(self.mz_exact), ptr(self.mz_pow), ptr(self.int_pow), ptr(self.norm),
                         self.max_peaks, float(m))


class SpectraSlice:
    """Spectra [start, stop) of a DeviceSpectra: the C struct points at the
    parent's offsets/norm shifted by ``start``; peak arrays are shared."""

    def __init__(self, parent: DeviceSpectra, start: int, stop: int):
        self.parent, self.start, self.stop = parent, start, stop
        self.n_spectra = stop - start
        lens = parent.packed.lengths[start:stop]
        self.max_peaks = int(lens.max()) if len(lens) else 0
        self.n_peaks = int(parent.packed.offsets[stop] - parent.packed.offsets[start])

    def struct(self):
        s = self.parent.struct()
        s.n_spectra = self.n_spectra
        s.n_peaks = self.n_peaks
        s.offsets = self.parent.offsets.data_ptr() + 8 * self.start
        s.norm = self.parent.norm.data_ptr() + 8 * self.start
        s.max_peaks = self.max_peaks
        return s
