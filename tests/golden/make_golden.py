"""Writes the golden fixtures under tests/golden/.

Run from the repository root: ``python tests/golden/make_golden.py``.

Fixture sources:
* cosine_greedy_kats.json -- inputs and expected matched index lists of the
  four known-answer cases in the reference's
  tests/similarity/test_cosine_greedy.py:27-95 and the 2 x 2 matrix case at
  :219-249 (transcribed data).
* fingerprint_kats.json -- fingerprints and expected values of
  tests/similarity/test_fingerprint_similarity.py:8-112 (transcribed data).
* pesticides.mgf -- the data file the reference's tests hold
  (tests/data/pesticides.mgf), copied byte for byte when the reference tree is
  present (the committed copy is used otherwise).
* pesticides_matchms.json -- matchms CosineGreedy top hits for query index 15
  (scan 2161) as printed in the reference notebook
  notebooks/samples/colab_tutorial_pesticide.ipynb (cell 10 output).
* pesticides_oracle.npz -- the oracle's 76 x 76 matrix (score, matches,
  candidates) for the MGF with intensities scaled to max 1, tol 0.1.
* uniform_slice.npz -- the oracle on the first 32 x 32 spectra of the
  256 x 256 uniform workload (refs seed 1, queries seed 2), tol 0.1.
"""
from __future__ import annotations

import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "cosine-similarity_amd"))

REFERENCE_MGF = "/root/reference/tests/data/pesticides.mgf"

COSINE_KATS = [
    {"ref_mz": [100, 200, 300, 500, 510], "ref_int": [0.1, 0.2, 1.0, 0.3, 0.4],
     "query_mz": [100, 200, 290, 490, 510], "query_int": [0.1, 0.2, 1.0, 0.3, 0.4],
     "tolerance": 0.1, "mz_power": 0.0, "intensity_power": 1.0,
     "expected_matches": [[0, 1, 4], [0, 1, 4]]},
    {"ref_mz": [100, 299, 300, 301, 510], "ref_int": [0.1, 1.0, 0.2, 0.3, 0.4],
     "query_mz": [100, 300, 301, 511], "query_int": [0.1, 1.0, 0.3, 0.4],
     "tolerance": 0.2, "mz_power": 0.0, "intensity_power": 1.0,
     "expected_matches": [[0, 2, 3], [0, 1, 2]]},
    {"ref_mz": [100, 299, 300, 301, 510], "ref_int": [0.1, 1.0, 0.2, 0.3, 0.4],
     "query_mz": [100, 300, 301, 511], "query_int": [0.1, 1.0, 0.3, 0.4],
     "tolerance": 2.0, "mz_power": 0.0, "intensity_power": 1.0,
     "expected_matches": [[0, 1, 3, 4], [0, 1, 2, 3]]},
    {"ref_mz": [100, 200, 300, 500, 510], "ref_int": [0.1, 0.2, 1.0, 0.3, 0.4],
     "query_mz": [100, 200, 290, 490, 510], "query_int": [0.1, 0.2, 1.0, 0.3, 0.4],
     "tolerance": 1.0, "mz_power": 0.5, "intensity_power": 2.0,
     "expected_matches": [[0, 1, 4], [0, 1, 4]]},
]

COSINE_MATRIX_KAT = {
    "spectra": [
        {"mz": [100, 200, 300], "int": [0.1, 0.2, 1.0]},
        {"mz": [110, 190, 290], "int": [0.5, 0.2, 1.0]},
    ],
    "rel": 1e-6,
}

FINGERPRINT_KATS = {
    "pair": {
        "fp1": [1, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0],
        "fp2": [0, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1],
        "expected": {"cosine": 0.6761234, "jaccard": 0.5, "dice": 2.0 / 3.0},
        "rel": 1e-6,
    },
    "empty": {
        "fp1": [0] * 16,
        "fp2": [0, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1],
        "expected_matrix": [[0, 0], [0, 1.0]],
        "rel": 1e-3,
    },
    "missing": {
        "fp1": [0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0],
        "fp2": [0, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1],
        "cases": [
            {"measure": "cosine", "expected": 0.84515425, "array_type": "numpy", "empty": "nan"},
            {"measure": "jaccard", "expected": 0.71428571, "array_type": "sparse", "empty": "nan"},
            {"measure": "dice", "expected": 0.83333333, "array_type": "numpy", "empty": 0},
        ],
    },
}

# notebooks/samples/colab_tutorial_pesticide.ipynb, cell 10 (matchms CosineGreedy,
# default_filters + normalize_intensities, query spectrums[15] = scan 2161)
PESTICIDE_MATCHMS = {
    "query_index": 15,
    "query_scan": "2161",
    "tolerance": 0.1,
    "hits": [
        {"scan": "613", "score": 0.8646, "matches": 14},
        {"scan": "603", "score": 0.8237, "matches": 14},
        {"scan": "2160", "score": 0.8015, "matches": 25},
        {"scan": "2362", "score": 0.2923, "matches": 7},
        {"scan": "2598", "score": 0.2231, "matches": 5},
        {"scan": "2594", "score": 0.1761, "matches": 3},
        {"scan": "1944", "score": 0.1396, "matches": 5},
        {"scan": "1772", "score": 0.1037, "matches": 4},
        {"scan": "2284", "score": 0.0675, "matches": 9},
    ],
}


def main() -> None:
    from oracle import oracle
    from cudams.spectrum import load_from_mgf, normalize_intensities
    from cudams import synthetic

    with open(os.path.join(HERE, "cosine_greedy_kats.json"), "w") as fh:
        json.dump({"pairs": COSINE_KATS, "matrix": COSINE_MATRIX_KAT}, fh, indent=1)
    with open(os.path.join(HERE, "fingerprint_kats.json"), "w") as fh:
        json.dump(FINGERPRINT_KATS, fh, indent=1)
    with open(os.path.join(HERE, "pesticides_matchms.json"), "w") as fh:
        json.dump(PESTICIDE_MATCHMS, fh, indent=1)

    mgf = os.path.join(HERE, "pesticides.mgf")
    if os.path.exists(REFERENCE_MGF):
        shutil.copyfile(REFERENCE_MGF, mgf)
    spectra = [normalize_intensities(s) for s in load_from_mgf(mgf)]
    pairs = [(s.peaks.mz, s.peaks.intensities) for s in spectra]
    res = oracle.cosine_greedy_matrix(pairs, pairs, tolerance=0.1)
    np.savez_compressed(os.path.join(HERE, "pesticides_oracle.npz"), score=res["score"],
                        matches=res["matches"], candidates=res["n_candidates"])

    refs = synthetic.uniform_spectra(256, 1)
    ques = synthetic.uniform_spectra(256, 2)
    r = [(refs.mz[refs.offsets[i]:refs.offsets[i + 1]],
          refs.intensities[refs.offsets[i]:refs.offsets[i + 1]]) for i in range(32)]
    q = [(ques.mz[ques.offsets[i]:ques.offsets[i + 1]],
          ques.intensities[ques.offsets[i]:ques.offsets[i + 1]]) for i in range(32)]
    res = oracle.cosine_greedy_matrix(r, q, tolerance=0.1)
    np.savez_compressed(os.path.join(HERE, "uniform_slice.npz"), score=res["score"],
                        matches=res["matches"], candidates=res["n_candidates"],
                        ref_mz=np.concatenate([x[0] for x in r]),
                        query_mz=np.concatenate([x[0] for x in q]))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
