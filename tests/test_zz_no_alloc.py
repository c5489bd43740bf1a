"""cs_cosine_greedy_noalloc (the workspace-free variant, its workspace a
file-scope device array) against cs_cosine_greedy with a caller workspace, on
two calls in a row, the second with a smaller input: bit-identical score,
matches and overflow.  Also: a query set too large for the fixed array is
refused, and the first workload computed again afterwards still agrees."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_both(refs, queries, tol):
    import torch

    from cudams.device import cosine_greedy, cosine_greedy_noalloc
    from cudams.spectra import DeviceSpectra

    R, Q = DeviceSpectra(refs), DeviceSpectra(queries)
    shape = (R.n_spectra, Q.n_spectra)
    outs = []
    for fn in ("ws", "noalloc"):
        score = torch.full(shape, -7.0, dtype=torch.float32, device="cuda")
        matches = torch.full(shape, -7, dtype=torch.int32, device="cuda")
        overflow = torch.full(shape, 7, dtype=torch.uint8, device="cuda")
        if fn == "ws":
            cosine_greedy(R, Q, tol, 0.0, 2048, score, matches, overflow)
        else:
            cosine_greedy_noalloc(R, Q, tol, 0.0, 2048, score, matches, overflow)
        torch.cuda.synchronize()
        outs.append((score.cpu().numpy(), matches.cpu().numpy(), overflow.cpu().numpy()))
    return outs


def assert_same(outs):
    (s0, m0, o0), (s1, m1, o1) = outs
    np.testing.assert_array_equal(m1, m0)
    np.testing.assert_array_equal(o1, o0)
    assert np.array_equal(s1.view(np.uint32), s0.view(np.uint32))
    assert (m0 >= 0).all()


def test_noalloc_equals_workspace_call_twice(gpu):
    from cudams import synthetic

    big = synthetic.lognormal_spectra(1024, 5)
    print("noalloc 1024x1024 clustered", flush=True)
    first = run_both(big, big, 0.01)
    assert_same(first)
    assert first[0][1].max() > 64  # heavy pairs went through the shared queue
    small_r, small_q = synthetic.uniform_spectra(100, 1), synthetic.uniform_spectra(70, 2)
    print("noalloc 100x70 uniform", flush=True)
    assert_same(run_both(small_r, small_q, 0.1))
    print("noalloc 1024x1024 again", flush=True)
    again = run_both(big, big, 0.01)
    assert_same(again)
    np.testing.assert_array_equal(again[1][1], first[1][1])


def test_noalloc_rejects_oversized_queries(gpu):
    """Refused in the size check, before any launch (real buffers all the same)."""
    import torch

    from cudams import _native, synthetic
    from cudams.spectra import DeviceSpectra

    L = _native.load()
    q = DeviceSpectra(synthetic.uniform_spectra(8, 3))
    qs = q.struct()
    qs.n_peaks = 1 << 23  # about 100 MB of index: past the fixed array
    rs = q.struct()
    p = _native.CsCosineParams(0.1, 0.0, 2048, 0, 0, 0, 0)
    score = torch.zeros((8, 8), dtype=torch.float32, device="cuda")
    matches = torch.zeros((8, 8), dtype=torch.int32, device="cuda")
    overflow = torch.zeros((8, 8), dtype=torch.uint8, device="cuda")
    st = L.cs_cosine_greedy_noalloc(ctypes.byref(rs), ctypes.byref(qs), ctypes.byref(p),
                                    score.data_ptr(), matches.data_ptr(), overflow.data_ptr(), 8,
                                    None)
    assert st == _native.CS_ERR_UNSUPPORTED
    assert b"allocation-free" in L.cs_last_error()
