"""Batches of independent CosineGreedy matrices (cs_cosine_batch_*): every
job equals its own cs_cosine_greedy call bit for bit (same kernels, same
arithmetic) and the oracle within the usual bar; graph replays equal the
eager run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCORE_ATOL = 1e-5


def say(msg):
    print(msg, flush=True)


def make_jobs(seed, n_jobs, mz_power, clusters=0.0, max_r=300, max_q=600):
    from cudams import synthetic
    from cudams.batch import CosineJob

    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_jobs):
        R = int(np.exp(rng.uniform(np.log(16), np.log(max_r))))
        Q = int(np.exp(rng.uniform(np.log(16), np.log(max_q))))
        tol = (0.01, 0.1)[k % 2]
        s = int(rng.integers(1 << 30))
        jobs.append(CosineJob(synthetic.lognormal_spectra(R, s, cluster_fraction=clusters),
                              synthetic.lognormal_spectra(Q, s + 1, cluster_fraction=clusters),
                              tol, 0.0, mz_power, 1.0, 2048))
    return jobs


def run_both(jobs, params=None):
    import torch

    from cudams.batch import BatchParams, BatchRunner, _DeviceJob, per_job

    runner = BatchRunner(jobs, params or BatchParams())
    runner.run()
    got = runner.results()
    ref_jobs = [_DeviceJob(j, torch.device("cuda")) for j in jobs]
    per_job(ref_jobs)
    torch.cuda.synchronize()
    want = [d.result() for d in ref_jobs]
    runner.close()
    return got, want


@pytest.mark.parametrize("mz_power", [0.0, 0.5])
def test_batch_equals_one_call_per_job(gpu, oracle, mz_power):
    jobs = make_jobs(61, 14, mz_power, clusters=0.2)
    say(f"batch of {len(jobs)} jobs, mz_power {mz_power}")
    got, want = run_both(jobs)
    for k, (g, w) in enumerate(zip(got, want)):
        for a, b in zip(g, w):
            np.testing.assert_array_equal(a, b, err_msg=f"job {k}")
    # the smallest jobs against the oracle
    order = sorted(range(len(jobs)), key=lambda k: jobs[k].refs.n_spectra * jobs[k].queries.n_spectra)
    for k in order[:3]:
        j = jobs[k]
        o = oracle.cosine_greedy_matrix((j.refs.mz, j.refs.intensities, j.refs.offsets),
                                        (j.queries.mz, j.queries.intensities, j.queries.offsets),
                                        tolerance=j.tolerance, mz_power=mz_power)
        np.testing.assert_array_equal(got[k][1], o["matches"])
        np.testing.assert_allclose(got[k][0], o["score"], rtol=0, atol=SCORE_ATOL)
        np.testing.assert_array_equal(got[k][2], (o["n_candidates"] >= 2048).astype(np.uint8))


def test_mixed_batch_variants(gpu, oracle):
    """Jobs that can take the float32 lane kernel (mz_power 0, float32 m/z, no
    shift) mixed with jobs that cannot (mz_power 0.5; float64 m/z): each class
    runs on its own variant in its own persistent launch.  Every job equals
    its own call bit for bit, and the smallest of each class the oracle."""
    from cudams import synthetic
    from cudams.batch import CosineJob
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(64)
    jobs = []
    for k in range(12):
        R = int(np.exp(rng.uniform(np.log(16), np.log(200))))
        Q = int(np.exp(rng.uniform(np.log(16), np.log(400))))
        s = int(rng.integers(1 << 30))
        refs = synthetic.lognormal_spectra(R, s, cluster_fraction=0.2)
        ques = synthetic.lognormal_spectra(Q, s + 1, cluster_fraction=0.2)
        if k % 4 == 3:  # m/z that are not float32 numbers: the float64 path
            ques = PackedSpectra(ques.mz + 1e-7, ques.intensities, ques.offsets)
        jobs.append(CosineJob(refs, ques, (0.01, 0.1)[k % 2], 0.0, 0.5 if k % 4 == 1 else 0.0,
                              1.0, 2048))
    say(f"mixed batch of {len(jobs)} jobs")
    got, want = run_both(jobs)
    for k, (g, w) in enumerate(zip(got, want)):
        for a, b in zip(g, w):
            np.testing.assert_array_equal(a, b, err_msg=f"job {k}")
    size = lambda k: jobs[k].refs.n_spectra * jobs[k].queries.n_spectra
    for cls in (0, 1, 3):
        k = min((k for k in range(len(jobs)) if k % 4 == cls), key=size)
        j = jobs[k]
        o = oracle.cosine_greedy_matrix((j.refs.mz, j.refs.intensities, j.refs.offsets),
                                        (j.queries.mz, j.queries.intensities, j.queries.offsets),
                                        tolerance=j.tolerance, mz_power=j.mz_power)
        np.testing.assert_array_equal(got[k][1], o["matches"], err_msg=f"job {k}")
        np.testing.assert_allclose(got[k][0], o["score"], rtol=0, atol=SCORE_ATOL)


def test_batch_launch_shapes_agree(gpu):
    """Different work-item sizes, register candidates and grids give the same
    results (the shape only changes who computes a pair)."""
    from cudams.batch import BatchParams

    jobs = make_jobs(62, 9, 0.0, clusters=0.3)
    base, _ = run_both(jobs)
    for bp in (BatchParams(4, 64, 64, 7), BatchParams(8, 1024, 1024, 0), BatchParams(16, 4, 32, 3)):
        say(f"shape {bp}")
        got, _ = run_both(jobs, bp)
        for g, b in zip(got, base):
            for x, y in zip(g, b):
                np.testing.assert_array_equal(x, y)


def test_batch_with_giant_pairs(gpu, oracle):
    """A batch holding a job whose pairs have tens of thousands of candidates
    (round-based path) next to ordinary jobs."""
    from cudams.batch import CosineJob
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(77)
    mzs, its = [], []
    for x in range(6):
        k = int(rng.integers(60, 200))
        mzs.append(np.sort(300.0 + rng.uniform(0, 0.8, k).astype(np.float32)).astype(np.float64))
        its.append(np.round(rng.random(k) * 8 + 1) / 8)
    dense = PackedSpectra.from_arrays(mzs, its)
    jobs = make_jobs(63, 4, 0.0) + [CosineJob(dense, dense, 1.0, 0.0, 0.0, 1.0, 5000)]
    got, want = run_both(jobs)
    for g, w in zip(got, want):
        for a, b in zip(g, w):
            np.testing.assert_array_equal(a, b)
    o = oracle.cosine_greedy_matrix((dense.mz, dense.intensities, dense.offsets),
                                    (dense.mz, dense.intensities, dense.offsets), tolerance=1.0,
                                    match_limit=5000)
    assert o["n_candidates"].max() > 4096
    np.testing.assert_array_equal(got[-1][1], o["matches"])
    np.testing.assert_allclose(got[-1][0], o["score"], rtol=0, atol=SCORE_ATOL)


def test_batch_graph_replays_equal_eager(gpu):
    """Capture one run; replay it three times with eager work in between
    (outputs cleared by a fill kernel); every replay equals the eager run."""
    import torch

    from cudams.batch import BatchRunner

    jobs = make_jobs(64, 10, 0.5, clusters=0.2)
    runner = BatchRunner(jobs)
    runner.run()
    eager = runner.results()
    runner.capture()
    for rep in range(3):
        for d in runner.djobs:
            d.score.fill_(-3.0)
            d.matches.fill_(-3)
            d.overflow.fill_(3)
        torch.cuda.synchronize()
        runner.replay()
        got = runner.results()
        for g, e in zip(got, eager):
            for a, b in zip(g, e):
                np.testing.assert_array_equal(a, b, err_msg=f"replay {rep}")
    runner.close()


def test_batch_heavy_queue_overflow(gpu):
    """Two jobs whose pairs all have 70 candidates (2 x 740^2 > 2^20 queued
    pairs): the shared queue overflows and every job's matrix is scanned."""
    from cudams.batch import BatchRunner, CosineJob
    from cudams.spectra import PackedSpectra

    rng = np.random.default_rng(22)
    k = 70
    mz = 100.0 + np.arange(k, dtype=np.float64)
    jobs = []
    for n in (740, 741):
        its = rng.random((n, k)).astype(np.float32).astype(np.float64) + 0.05
        p = PackedSpectra.from_arrays([mz] * n, list(its))
        jobs.append(CosineJob(p, p, 0.1))
    say("batch heavy queue overflow")
    runner = BatchRunner(jobs)
    runner.run()
    for j, (s, m, o) in zip(jobs, runner.results()):
        n = j.refs.n_spectra
        w = j.refs.intensities.reshape(n, k)
        norm = np.sqrt((w * w).sum(axis=1))
        assert np.all(m == k)
        np.testing.assert_allclose(s, (w @ w.T) / np.outer(norm, norm), rtol=0, atol=SCORE_ATOL)
    runner.close()


def test_autotune_picks_a_timed_shape(gpu, tmp_path):
    from cudams.batch import autotune

    jobs = make_jobs(65, 6, 0.0)
    cache = str(tmp_path / "tune.json")
    best, timings = autotune(jobs, {"refs_per_block": (64, 256), "light_candidates": (8,)},
                             replays=2, cache_path=cache)
    assert best in timings and len(timings) == 2
    again, t2 = autotune(jobs, cache_path=cache)
    assert again == best and t2 == {}


def test_batched_jobs_workload_vs_oracle(gpu, oracle):
    """The 64-job workload at its stated size (seed 6: R x Q log-uniform in
    [16, 2048] x [16, 4096], tolerance 0.01 or 0.1, mz_power 0.5; 14,966,214
    pairs), captured in a HIP graph and replayed; every job of the replay
    against the oracle: matches and overflow exact, score within 1e-5."""
    import os as _os

    import torch

    from cudams import synthetic
    from cudams.batch import BatchRunner

    jobs = synthetic.batched_jobs(seed=6)
    assert len(jobs) == 64
    say("64 batched jobs: capture + replay")
    runner = BatchRunner(jobs)
    runner.run()
    runner.capture()
    for d in runner.djobs:
        d.matches.fill_(-3)
    torch.cuda.synchronize()
    runner.replay()
    got = runner.results()
    runner.close()
    cores = len(_os.sched_getaffinity(0))
    pairs = 0
    for k, (j, (s, m, o)) in enumerate(zip(jobs, got)):
        w = oracle.cosine_greedy_matrix((j.refs.mz, j.refs.intensities, j.refs.offsets),
                                        (j.queries.mz, j.queries.intensities, j.queries.offsets),
                                        tolerance=j.tolerance, mz_power=j.mz_power,
                                        intensity_power=j.intensity_power, nthreads=cores)
        np.testing.assert_array_equal(m, w["matches"], err_msg=f"job {k}")
        np.testing.assert_array_equal(o, (w["n_candidates"] >= j.match_limit).astype(np.uint8),
                                      err_msg=f"job {k}")
        np.testing.assert_allclose(s, w["score"], rtol=0, atol=SCORE_ATOL, err_msg=f"job {k}")
        pairs += m.size
    assert pairs == 14_966_214
