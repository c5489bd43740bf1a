"""A minimal spectrum type with the attributes the similarity classes read.

The reference takes matchms ``Spectrum`` objects (``.peaks.mz``,
``.peaks.intensities``, ``.get("fingerprint")``; CudaCosineGreedy.py:68,82-84,
CudaFingerprintSimilarity.py:98-99).  matchms is optional here: when it is
installed its spectra work unchanged; otherwise this class offers the same
attributes, and :func:`load_from_mgf` reads MGF files into it.
"""
from __future__ import annotations

from typing import Iterator, Optional

import numpy as np


class Fragments:
    """Peaks of one spectrum, m/z ascending (float64)."""

    def __init__(self, mz: np.ndarray, intensities: np.ndarray):
        mz = np.asarray(mz, dtype=np.float64)
        intensities = np.asarray(intensities, dtype=np.float64)
        if mz.shape != intensities.shape or mz.ndim != 1:
            raise ValueError("m/z and intensities must be 1-D arrays of equal length")
        if mz.size > 1 and np.any(np.diff(mz) < 0):
            raise ValueError("m/z values must be sorted ascending")
        self.mz = mz
        self.intensities = intensities

    def __len__(self) -> int:
        return len(self.mz)

    @property
    def to_numpy(self) -> np.ndarray:
        return np.stack([self.mz, self.intensities], axis=1)


class Spectrum:
    """Peaks plus a metadata dict (keys lower-cased, as matchms stores them)."""

    def __init__(self, mz: np.ndarray, intensities: np.ndarray, metadata: Optional[dict] = None,
                 metadata_harmonization: bool = False):
        self._peaks = Fragments(mz, intensities)
        self.metadata = {str(k).lower(): v for k, v in (metadata or {}).items()}

    @property
    def peaks(self) -> Fragments:
        return self._peaks

    def get(self, key: str, default=None):
        return self.metadata.get(key.lower(), default)

    def set(self, key: str, value) -> "Spectrum":
        self.metadata[key.lower()] = value
        return self

    def __len__(self) -> int:
        return len(self._peaks)

    def __repr__(self) -> str:
        return f"Spectrum({len(self)} peaks, {len(self.metadata)} metadata keys)"


def normalize_intensities(s: Spectrum) -> Spectrum:
    """Scale intensities to a maximum of 1 (matchms normalize_intensities)."""
    it = s.peaks.intensities
    if it.size == 0 or it.max() <= 0:
        return s
    out = Spectrum(s.peaks.mz.copy(), it / it.max())
    out.metadata = dict(s.metadata)
    return out


def load_from_mgf(path: str) -> Iterator[Spectrum]:
    """Read an MGF file: ``BEGIN IONS`` / ``KEY=value`` lines / ``mz intensity``
    peak lines / ``END IONS``.  Peaks are sorted by m/z (stable)."""
    meta, mz, it, inside = {}, [], [], False
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line[0] in "#;!/":
                continue
            if line == "BEGIN IONS":
                meta, mz, it, inside = {}, [], [], True
            elif line == "END IONS":
                if inside:
                    m = np.asarray(mz, np.float64)
                    i = np.asarray(it, np.float64)
                    o = np.argsort(m, kind="stable")
                    yield Spectrum(m[o], i[o], meta)
                inside = False
            elif inside:
                head = line[0]
                if head.isdigit() or head in "+-.":
                    parts = line.split()
                    mz.append(float(parts[0]))
                    it.append(float(parts[1]) if len(parts) > 1 else 1.0)
                elif "=" in line:
                    k, v = line.split("=", 1)
                    meta[k.strip().lower()] = v.strip()
