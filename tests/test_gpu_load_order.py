"""A process that loads the library before it imports torch: one
CosineGreedy call and one two-job batch must run and agree with the oracle.
torch's wheel carries its own HIP runtime; the loader imports torch before
the library so that the library binds to that runtime, and the process has
one.  (With the library loaded first, torch brought a second runtime, and
the library's then found no device.)  The order is set in a child process,
since this one has torch loaded already."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCORE_ATOL = 1e-5

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
from cudams import _native
_native.load()                       # the first import of the process
from cudams import synthetic
import torch
from cudams.batch import BatchRunner, CosineJob
from cudams.device import Workspace, cosine_greedy
from cudams.spectra import DeviceSpectra

dev = torch.device("cuda:0")
refs = synthetic.lognormal_spectra(37, 11, cluster_fraction=0.3)
ques = synthetic.lognormal_spectra(53, 12, cluster_fraction=0.3)
R, Q = DeviceSpectra(refs, 0.0, 1.0, dev), DeviceSpectra(ques, 0.0, 1.0, dev)
score = torch.empty((37, 53), dtype=torch.float32, device=dev)
matches = torch.empty((37, 53), dtype=torch.int32, device=dev)
overflow = torch.empty((37, 53), dtype=torch.uint8, device=dev)
cosine_greedy(R, Q, 0.01, 0.0, 2048, score, matches, overflow, Workspace(dev))
runner = BatchRunner([CosineJob(refs, ques, 0.01), CosineJob(ques, refs, 0.1)])
runner.run()
got = runner.results()
runner.close()
torch.cuda.synchronize()
print(json.dumps({"score": score.cpu().tolist(), "matches": matches.cpu().tolist(),
                  "batch": [[s.tolist(), m.tolist()] for s, m, o in got]}))
"""


def test_library_loaded_before_torch(gpu, oracle):
    from cudams import synthetic

    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "cosine-similarity_amd")
    print("child: library before torch, one call and a two-job batch", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, pkg], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    refs = synthetic.lognormal_spectra(37, 11, cluster_fraction=0.3)
    ques = synthetic.lognormal_spectra(53, 12, cluster_fraction=0.3)
    cases = [((refs, ques, 0.01), out["score"], out["matches"]),
             ((refs, ques, 0.01), *out["batch"][0]),
             ((ques, refs, 0.1), *out["batch"][1])]
    for (a, b, tol), s, m in cases:
        w = oracle.cosine_greedy_matrix((a.mz, a.intensities, a.offsets),
                                        (b.mz, b.intensities, b.offsets), tolerance=tol)
        np.testing.assert_array_equal(np.asarray(m), w["matches"])
        np.testing.assert_allclose(np.asarray(s, np.float64), w["score"], rtol=0, atol=SCORE_ATOL)
