"""Served library search on the GPU: the shared-index entry point against
the plain one, and request results that are the same alone and in the mix.
Bar: bit-exact between paths (same pairs, same arithmetic); against the
oracle, top-k scores within 1e-5 of its float64 scores."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCORE_ATOL = 1e-5


def say(msg):
    print(msg, flush=True)


def test_indexed_calls_share_one_index_on_two_streams(gpu):
    """One PREPARE of a query set; two reference sets searched concurrently on
    two streams with their own scratch; each equals cs_cosine_greedy."""
    import torch

    from cudams import _native, synthetic
    from cudams.device import (cosine_greedy, cosine_greedy_indexed,
                               cosine_greedy_indexed_bytes)
    from cudams.spectra import DeviceSpectra, PackedSpectra

    lib = DeviceSpectra(synthetic.lognormal_spectra(700, 31, cluster_fraction=0.3))
    ib, sb = cosine_greedy_indexed_bytes(lib)
    index = torch.empty(ib, dtype=torch.uint8, device="cuda")
    empty = DeviceSpectra(PackedSpectra.from_arrays([], []))
    cosine_greedy_indexed(empty, lib, 0.01, 0.0, 2048, None, None, None, index, None,
                          stages=_native.CS_STAGE_PREPARE)
    refsets = [DeviceSpectra(synthetic.lognormal_spectra(n, s, cluster_fraction=0.3))
               for n, s in ((300, 32), (45, 33))]
    outs, streams = [], [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for R, st in zip(refsets, streams):
        o = [torch.empty((R.n_spectra, 700), dtype=dt, device="cuda")
             for dt in (torch.float32, torch.int32, torch.uint8)]
        scratch = torch.empty(sb, dtype=torch.uint8, device="cuda")
        with torch.cuda.stream(st):
            cosine_greedy_indexed(R, lib, 0.01, 0.0, 2048, *o, index, scratch,
                                  stages=_native.CS_STAGE_LIGHT | _native.CS_STAGE_HEAVY,
                                  stream=st)
        outs.append((o, scratch))
    torch.cuda.synchronize()
    for R, (o, _) in zip(refsets, outs):
        w = [torch.empty((R.n_spectra, 700), dtype=dt, device="cuda")
             for dt in (torch.float32, torch.int32, torch.uint8)]
        cosine_greedy(R, lib, 0.01, 0.0, 2048, *w)
        torch.cuda.synchronize()
        for a, b in zip(o, w):
            assert torch.equal(a, b)


def _library_and_requests(n_lib=6000, n_req=40, seed=41):
    from cudams import synthetic
    from cudams.serving import pack_request
    from cudams.spectra import PackedSpectra

    lib = synthetic.lognormal_spectra(n_lib, seed, cluster_fraction=0.1)
    pool = synthetic.lognormal_spectra(600, seed + 1, cluster_fraction=0.1)
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        size = int(rng.integers(1, 40))
        a = int(rng.integers(0, pool.n_spectra - size))
        off = pool.offsets[a:a + size + 1]
        p = PackedSpectra(pool.mz[off[0]:off[-1]], pool.intensities[off[0]:off[-1]],
                          off - off[0])
        reqs.append((p, pack_request(p)))
    return lib, reqs


def test_served_results_same_alone_and_in_the_mix(gpu, oracle):
    """Every request searched alone (plain per-request call) and fused with
    the others by the continuous-batching server (two slots, a budget that
    forces several batches) gives identical hits; a few rows are checked
    against the oracle."""
    from cudams.serving import LibraryIndex, PlainSearch, SearchServer, pack_request

    lib, reqs = _library_and_requests()
    ix = LibraryIndex(lib, tolerance=0.01)
    plain = PlainSearch(ix, k=10)
    say("alone")
    alone = [plain.search(r) for _, r in reqs]
    say("in the mix")
    budget = 2 * 9 * ix.n * 64  # two slots of 64 request spectra
    with SearchServer(ix, k=10, slots=2, memory_budget=budget) as srv:
        futs = [srv.submit(pack_request(p)) for p, _ in reqs]
        mixed = [f.result(timeout=120) for f in futs]
    assert srv.batches < len(reqs)  # requests were fused
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a.indices, b.indices)
        np.testing.assert_array_equal(a.scores, b.scores)
        np.testing.assert_array_equal(a.matches, b.matches)
    # oracle: the returned scores are the oracle's scores of those library
    # spectra, and no library spectrum outside the list scores higher
    csr = (lib.mz, lib.intensities, lib.offsets)
    for (p, _), hits in list(zip(reqs, alone))[:3]:
        want = oracle.cosine_greedy_matrix((p.mz, p.intensities, p.offsets), csr, tolerance=0.01)
        for row in range(p.n_spectra):
            s = want["score"][row]
            idx = hits.indices[row]
            np.testing.assert_allclose(hits.scores[row], s[idx], rtol=0, atol=SCORE_ATOL)
            np.testing.assert_array_equal(hits.matches[row], want["matches"][row][idx])
            kth = hits.scores[row][-1]
            outside = np.delete(s, idx)
            assert np.all(outside <= kth + SCORE_ATOL)


def test_server_rejects_oversized_request(gpu):
    from cudams.serving import LibraryIndex, SearchServer
    from cudams import synthetic

    lib = synthetic.lognormal_spectra(500, 3, cluster_fraction=0.0)
    ix = LibraryIndex(lib)
    srv = SearchServer(ix, k=5, slots=1, memory_budget=9 * ix.n * 4)
    big = synthetic.lognormal_spectra(5, 4, cluster_fraction=0.0)
    with pytest.raises(ValueError):
        srv.submit(big)


def test_served_library_workload_vs_oracle(gpu, oracle):
    """The served workload at its stated size: a 500,000-spectrum library
    (lognormal peaks, seed 7) indexed once; requests of 1, 16 and 64 spectra
    (query pool seed 8) served together through the continuous-batching
    server and one by one through the plain path.  Both give identical hits;
    against the oracle over the whole library, each row's ten scores are the
    ten best within 1e-5 (no library spectrum outside the list scores higher),
    their matches exact, and in order (score desc, library index asc)."""
    import os as _os

    from cudams import synthetic
    from cudams.serving import LibraryIndex, PlainSearch, SearchServer, pack_request
    from cudams.spectra import PackedSpectra

    lib = synthetic.lognormal_spectra(500_000, 7, cluster_fraction=0.0)
    pool = synthetic.lognormal_spectra(4096, 8, cluster_fraction=0.0)
    say("500k library: index")
    ix = LibraryIndex(lib, tolerance=0.01)
    reqs = []
    for start, size in ((0, 1), (100, 16), (1000, 64)):
        off = pool.offsets[start:start + size + 1]
        reqs.append(PackedSpectra(pool.mz[off[0]:off[-1]], pool.intensities[off[0]:off[-1]],
                                  off - off[0]))
    say("served")
    with SearchServer(ix, k=10, slots=3) as srv:
        futs = [srv.submit(pack_request(p)) for p in reqs]
        served = [f.result(timeout=300) for f in futs]
    say("plain")
    plain = PlainSearch(ix, k=10)
    alone = [plain.search(pack_request(p)) for p in reqs]
    csr = (lib.mz, lib.intensities, lib.offsets)
    cores = len(_os.sched_getaffinity(0))
    for p, a, b in zip(reqs, alone, served):
        np.testing.assert_array_equal(a.indices, b.indices)
        np.testing.assert_array_equal(a.scores, b.scores)
        np.testing.assert_array_equal(a.matches, b.matches)
        say(f"oracle: {p.n_spectra} x 500000")
        want = oracle.cosine_greedy_matrix((p.mz, p.intensities, p.offsets), csr,
                                           tolerance=0.01, nthreads=cores)
        for row in range(p.n_spectra):
            s, idx = want["score"][row], a.indices[row]
            assert (idx >= 0).all()
            np.testing.assert_allclose(a.scores[row], s[idx], rtol=0, atol=SCORE_ATOL)
            np.testing.assert_array_equal(a.matches[row], want["matches"][row][idx])
            best = np.sort(s)[::-1][:10]
            np.testing.assert_allclose(a.scores[row], best, rtol=0, atol=SCORE_ATOL)
            assert np.all(np.delete(s, idx) <= a.scores[row][-1] + SCORE_ATOL)
            order = np.lexsort((idx, -a.scores[row]))
            assert np.array_equal(order, np.arange(10))
