#!/usr/bin/env python3
"""Benchmark of the all-pairs CosineGreedy path on one MI355X.

Default workload: 8192 x 8192 spectra (refs = queries, as the reference's own
benchmark runs them), peaks round(lognormal(ln 60, 0.8)) clipped to
[5, 1024], 20% of spectra with a dense cluster of 30-80 peaks inside 0.05 Da,
m/z U[50, 1000] float32, intensities U(0, 1] scaled to max 1, tolerance 0.01,
mz_power 0, intensity_power 1, match_limit 2048.  Inputs are resident in HBM
before the timed region; one step is the work of one full cs_cosine_greedy
call (query tiling + lane kernel + cooperative kernel), issued as its three
stages with HIP events between them, producing score/matches/overflow for all
67,108,864 pairs.

Prints one JSON line (rank 0).  Beside the step it carries: ``roofline`` of
the lane kernel (compulsory bytes against HBM, VALU busy from the committed
counter profile), ``cpu_baseline`` (the oracle over the whole matrix on the
host cores), ``parity`` (every pair against that oracle, near ties counted),
``matrix_wall`` (the drop-in ``CudaCosineGreedy.matrix`` call end to end),
``cpu_parallel`` (the host engine behind ``CPUParallelCosineGreedy``),
``fingerprint`` (the three measures at 4096 x 4096 x 2048 bits),
``batched_jobs`` (64 uneven jobs, graph vs one call per job) and ``serving``
(500,000-spectrum library, Poisson request streams, served vs plain).
``--workload fingerprint`` runs the fingerprint section as its own line, and
``--workload batched`` the 64-job batch's captured graph alone (a fixed launch
shape, the one the autotuner picks in the default run; for profiling).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "cosine-similarity_amd"))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP4_PEAK_TFLOPS = 10100.0  # MI355X dense fp4 MFMA
INT8_PEAK_TOPS = 5000.0    # MI355X dense int8 MFMA


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters (see
    profiles/), or None when no measurement for this kernel/workload exists."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    e = t.get(f"{kernel}@{workload}")
    return None if e is None else e.get("hbm_bytes_per_launch")


def host_threads() -> int:
    """Host threads for the CPU legs: the cores this process may run on, capped
    by OMP_NUM_THREADS when the environment sets it (a GPU box allots each GPU
    a share of its cores that way)."""
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return cores


def cpu_baseline_cosine(packed, tol):
    """The oracle (C restatement of matchms CosineGreedy, float64, OpenMP over
    rows) on the whole matrix, on this host's cores.  Returns the baseline
    record and the oracle's matrix (used for the parity fields)."""
    from oracle import oracle

    cores = host_threads()
    n = packed.n_spectra
    csr = (packed.mz, packed.intensities, packed.offsets)
    t0 = time.perf_counter()
    want = oracle.cosine_greedy_matrix(csr, csr, tolerance=tol, nthreads=cores)
    t = time.perf_counter() - t0
    # one thread: the per-core rate, on a stripe of rows sized to about 2 s
    rows1 = max(1, min(n, int(n * n / t / cores * 2.0 / n)))
    sub1 = (packed.mz[: packed.offsets[rows1]], packed.intensities[: packed.offsets[rows1]],
            packed.offsets[: rows1 + 1])
    t1 = time.perf_counter()
    oracle.cosine_greedy_matrix(sub1, csr, tolerance=tol, nthreads=1)
    t1 = time.perf_counter() - t1
    rec = {"value": n * n / t, "unit": "pairs/s", "cores": cores, "kind": "port",
           "host_cores_visible": len(os.sched_getaffinity(0)),
           "sample": f"all {n} x {n} pairs ({n * n}), {t:.1f} s, oracle/cs_oracle.c float64, "
                     f"OpenMP over reference rows",
           "single_core": {"value": rows1 * n / t1, "unit": "pairs/s", "cores": 1,
                           "sample": f"first {rows1} reference rows x all {n} queries, "
                                     f"{t1:.1f} s"}}
    return rec, want


def cosine_parity(got: dict, want: dict, packed, tol, match_limit) -> dict:
    """GPU matrix against the oracle's: mismatch counts, the largest score
    error, and near ties.  ``near_tie_pairs`` counts the pairs whose accepted
    peak pairs change when candidate values are rounded to float32 (the
    reference GPU kernel's stored products); ``near_tie_reassignments`` counts
    the pairs where the GPU result departs from the float64 oracle: matches
    differ, or the score is more than two float32 ulps from the oracle's."""
    import numpy as np

    from oracle import oracle

    err = np.abs(got["score"].astype(np.float64) - want["score"])
    ulp = np.spacing(np.abs(want["score"]).astype(np.float32)).astype(np.float64)
    mism = got["matches"] != want["matches"]
    reassigned = mism | (err > 2.0 * ulp)
    csr = (packed.mz, packed.intensities, packed.offsets)
    ties = oracle.near_tie_pairs(csr, csr, tolerance=tol, nthreads=host_threads()).astype(bool)
    return {"pairs": int(err.size), "oracle": "oracle/cs_oracle.c float64, every pair",
            "matches_mismatch": int(mism.sum()),
            "overflow_mismatch": int((got["overflow"] != (want["n_candidates"] >= match_limit)).sum()),
            "max_abs_score_err": float(err.max()), "score_atol": 1e-5,
            "near_tie_pairs": int(ties.sum()),
            "near_tie_reassignments": int(reassigned.sum()),
            "near_tie_reassignments_among_near_tie_pairs": int((reassigned & ties).sum()),
            "overflow_pairs": int(want["overflow"].sum())}


def light_roofline(n, lens, light_ms) -> dict:
    """Roofline of the lane kernel (the LIGHT stage) against HBM.  Algorithmic bytes
    per launch are the compulsory ones: the 9 output bytes of every pair
    (score f32 + matches i32 + overflow u8) and each spectrum's peaks read
    once (float32 m/z + float32 intensity, 8 B per peak per side).  The
    streaming model that re-reads both spectra for every pair,
    8 (n_r + n_q) + 9 bytes per pair, is kept as ``effective_gbs``."""
    peaks = float(lens.sum())
    pairs = float(n) * n
    compulsory = 9.0 * pairs + 8.0 * (peaks + peaks)
    streaming = 8.0 * (n * peaks + n * peaks) + 9.0 * pairs
    t = light_ms / 1e3
    achieved = compulsory / t / 1e9
    rec = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS,
           "traffic": traffic_from_profiles("k_light", f"cosine_{n}"),
           "kernel": profile_field("k_light", f"cosine_{n}", "kernel_name") or "k_light",
           "kernel_ms": light_ms, "algorithmic_bytes": compulsory,
           "effective_gbs": streaming / t / 1e9,
           "effective_note": "8 (n_r + n_q) + 9 bytes per pair, every spectrum re-read per pair"}
    valu = profile_field("k_light", f"cosine_{n}", "valu_busy")
    if valu is not None:
        rec["limiter"] = "valu_issue"
        rec["valu_frac"] = valu
        rec["wait_frac"] = profile_field("k_light", f"cosine_{n}", "wave_wait")
    return rec


def profile_field(kernel: str, workload: str, field: str):
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        e = json.load(fh).get(f"{kernel}@{workload}")
    return None if e is None else e.get(field)


def bench_cosine(a) -> dict:
    import numpy as np
    import torch

    from cudams import _native, synthetic
    from cudams.device import Workspace, cosine_greedy
    from cudams.spectra import DeviceSpectra

    n, tol, ml = a.n, 0.01, 2048
    log(f"generating {n} clustered lognormal spectra")
    packed = synthetic.lognormal_spectra(n, 5)
    dev = torch.device("cuda:0")
    S = DeviceSpectra(packed, 0.0, 1.0, dev)
    score = torch.empty((n, n), dtype=torch.float32, device=dev)
    matches = torch.empty((n, n), dtype=torch.int32, device=dev)
    overflow = torch.empty((n, n), dtype=torch.uint8, device=dev)
    ws = Workspace(dev)

    def step(stages=0):
        cosine_greedy(S, S, tol, 0.0, ml, score, matches, overflow, ws, a.light_candidates,
                      a.round_capacity, stages=stages)

    log(f"warmup {a.warmup}")
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    # One timed step = the three stages of one cs_cosine_greedy call issued as
    # three calls on the same stream, with HIP events between them, so the
    # per-stage device time (k_light alone is the LIGHT stage) comes from the
    # timed steps themselves.
    stream = torch.cuda.current_stream()
    order = ((_native.CS_STAGE_PREPARE, "prepare"), (_native.CS_STAGE_LIGHT, "light"),
             (_native.CS_STAGE_HEAVY, "heavy"))
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(len(order) + 1)]
              for _ in range(a.steps)]
    log(f"timing {a.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ev = events[k]
        ev[0].record(stream)
        for x, (st, _) in enumerate(order):
            step(st)
            ev[x + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / a.steps
    stage_ms = {name: sum(events[k][x].elapsed_time(events[k][x + 1]) for k in range(a.steps)) / a.steps
                for x, (_, name) in enumerate(order)}
    log(f"stages ms {stage_ms}")

    lens = np.diff(packed.offsets).astype(np.float64)
    pairs = n * n
    res = {
        "metric": "CosineGreedy pairs/s at 8k x 8k; p50/p99 search latency served vs plain",
        "value": pairs / (ms / 1e3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cosine_greedy_8192x8192_lognormal_clusters", "refs": n,
                   "queries": n, "tolerance": tol, "mz_power": 0.0, "intensity_power": 1.0,
                   "match_limit": ml, "peaks": "round(lognormal(ln 60, 0.8)) in [5, 1024]",
                   "cluster_fraction": 0.2, "seed": 5, "mean_peaks": float(lens.mean())},
        "stage_ms": stage_ms,
        "roofline": light_roofline(n, lens, stage_ms["light"]),
    }
    got = {"score": score.cpu().numpy(), "matches": matches.cpu().numpy(),
           "overflow": overflow.cpu().numpy()}
    del score, matches, overflow
    want = None
    if not a.no_cpu_baseline:
        log("cpu baseline (the oracle over the whole matrix)")
        res["cpu_baseline"], want = cpu_baseline_cosine(packed, tol)
    else:
        res["cpu_baseline"] = None
    if not a.no_parity:
        if want is None:
            from oracle import oracle

            csr = (packed.mz, packed.intensities, packed.offsets)
            want = oracle.cosine_greedy_matrix(csr, csr, tolerance=tol, nthreads=host_threads())
        log("parity against the oracle")
        res["parity"] = cosine_parity(got, want, packed, tol, ml)
        log(f"parity {res['parity']}")
    if not a.no_cpu_baseline and not a.no_cpu_parallel:
        res["cpu_parallel"] = bench_cpu_parallel(packed, tol, want)
    del got, want
    if not a.no_matrix_wall:
        res["matrix_wall"] = bench_matrix_wall(a, packed)
    return res


def bench_fingerprint_measures(a) -> dict:
    """Jaccard, Dice and cosine on 4096 x 4096 fingerprints of 2048 bits
    (seeds 3 and 4, per-row density U[0.02, 0.3]).  Per measure: device time
    per call from HIP events on the launch stream, the matrix-core roofline
    (2 R Q K bit operations per call against the dense fp4 peak the kernel
    runs at, and against the int8 peak), the oracle on the host cores, and
    parity against it over the whole matrix."""
    import numpy as np
    import torch

    from cudams import synthetic
    from cudams.device import Workspace, fingerprint_similarity

    n, bits = 4096, 2048
    dev = torch.device("cuda:0")
    fa = synthetic.fingerprints(n, 3, n_bits=bits)
    fb = synthetic.fingerprints(n, 4, n_bits=bits)
    wa = torch.from_numpy(synthetic.pack_bits(fa).view(np.int32)).to(dev)
    wb = torch.from_numpy(synthetic.pack_bits(fb).view(np.int32)).to(dev)
    out = torch.empty((n, n), dtype=torch.float32, device=dev)
    ws = Workspace(dev)
    stream = torch.cuda.current_stream()
    ops = 2.0 * n * n * bits
    res = {"config": {"workload": "fingerprint_4096x4096x2048", "refs": n, "queries": n,
                      "bits": bits, "density": "per-row U[0.02, 0.3]", "seeds": [3, 4]},
           "dtype": "fp4 (0/1 bit planes, exact float32 counts)", "measures": {}}
    steps = max(a.steps, 10)
    for measure in ("jaccard", "dice", "cosine"):
        for _ in range(max(a.warmup, 1)):
            fingerprint_similarity(measure, wa, wb, out, workspace=ws)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fingerprint_similarity(measure, wa, wb, out, workspace=ws)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        tflops = ops / (ms / 1e3) / 1e12
        m = {"ms_per_call": ms, "pairs_per_s": n * n / (ms / 1e3),
             "roofline": {"bound": "mfma", "achieved": tflops, "peak": FP4_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": tflops / FP4_PEAK_TFLOPS,
                          "int8_frac": tflops / INT8_PEAK_TOPS, "traffic": None,
                          "output_hbm_frac": 4.0 * n * n / (ms / 1e3) / (HBM_PEAK_GBS * 1e9)}}
        got = out.cpu().numpy()
        if not a.no_cpu_baseline:
            from oracle import oracle

            cores = host_threads()
            t0 = time.perf_counter()
            want, _ = oracle.fingerprint_matrix(fa, fb, measure, nthreads=cores)
            t = time.perf_counter() - t0
            m["cpu_baseline"] = {"value": n * n / t, "unit": "pairs/s", "cores": cores,
                                 "kind": "port",
                                 "sample": f"all {n} x {n} pairs, {t:.1f} s, oracle/cs_oracle.c "
                                           f"fingerprint_matrix ({measure}), float64, OpenMP"}
            m["parity"] = {"pairs": int(got.size),
                           "equal_to_float32_of_oracle": int((got == want.astype(np.float32)).sum()),
                           "max_abs_err": float(np.abs(got.astype(np.float64) - want).max())}
        log(f"fingerprint {measure}: {ms * 1e3:.1f} us/call, {m['roofline']['frac']:.3f} of fp4")
        res["measures"][measure] = m
    return res


def bench_fingerprint(a) -> dict:
    """``--workload fingerprint``: the same section as its own bench line."""
    fp = bench_fingerprint_measures(a)
    j = fp["measures"]["jaccard"]
    return {"metric": "FingerprintSimilarity jaccard pairs/s at 4096 x 4096 x 2048 bits",
            "value": j["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
            "steps": max(a.steps, 10), "warmup": max(a.warmup, 1), "ms_per_step": j["ms_per_call"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp4",
            "data": "synthetic", "config": fp["config"], "roofline": j["roofline"],
            "cpu_baseline": j.get("cpu_baseline"), "fingerprint": fp}


def bench_batched_graph(a) -> dict:
    """``--workload batched``: replays of the 64-job batch's captured graph
    with the launch shape the default run's autotuner picks (K 8, round
    capacity 256, 512 references per block), nothing else on the stream, so
    a kernel trace of this run holds the batch kernels only."""
    import torch

    from cudams import synthetic
    from cudams.batch import BatchParams, BatchRunner

    jobs = synthetic.batched_jobs(seed=6)
    pairs = sum(j.refs.n_spectra * j.queries.n_spectra for j in jobs)
    runner = BatchRunner(jobs, BatchParams(8, 256, 512, 0))
    runner.capture()
    for _ in range(max(1, a.warmup)):
        runner.replay()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(a.steps):
        runner.replay()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    runner.close()
    log(f"batched graph: {ms:.3f} ms per replay, {pairs} pairs")
    return {"metric": "batched CosineGreedy pairs/s (64 uneven jobs, one graph)",
            "value": pairs / (ms / 1e3), "unit": "pairs/s", "n_gpus": 1, "steps": a.steps,
            "warmup": max(1, a.warmup), "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "batched_jobs_64_seed6", "jobs": len(jobs),
                                            "pairs": pairs, "mz_power": 0.5,
                                            "tolerances": [0.01, 0.1]},
            "roofline": None, "cpu_baseline": None}


def bench_cpu_parallel(packed, tol, want) -> dict:
    """CPUParallelCosineGreedy (the host engine, libcsamd_cpu.so) on the whole
    matrix with this host's cores, wall clock from spectrum objects to the
    structured array, and its agreement with the oracle when that matrix is
    at hand.  The reference keeps its CPU class for comparison with the GPU
    one (spectrum_similarity_functions.py:290-293)."""
    import numpy as np

    from cudams.similarity import CPUParallelCosineGreedy
    from cudams.spectrum import Spectrum

    off = packed.offsets
    spectra = [Spectrum(packed.mz[off[i]:off[i + 1]], packed.intensities[off[i]:off[i + 1]])
               for i in range(packed.n_spectra)]
    cores = host_threads()
    cpu = CPUParallelCosineGreedy(tolerance=tol, n_threads=cores)
    t0 = time.perf_counter()
    out = cpu.matrix(spectra, spectra)
    t = time.perf_counter() - t0
    n = packed.n_spectra
    rec = {"what": f"CPUParallelCosineGreedy(tolerance={tol}).matrix(spectra, spectra)",
           "value": n * n / t, "unit": "pairs/s", "cores": cores, "seconds": t}
    if want is not None:
        rec["matches_mismatch"] = int((out["matches"] != want["matches"]).sum())
        rec["max_abs_score_err"] = float(
            np.abs(out["score"].astype(np.float64) - want["score"]).max())
    log(f"cpu engine {t:.2f} s on {cores} threads: {rec}")
    return rec


def bench_matrix_wall(a, packed) -> dict:
    """What a user of the drop-in class waits for: CudaCosineGreedy.matrix on
    lists of spectrum objects, wall clock (packing, upload, kernels, download
    and the structured array), as the reference's own benchmark times it
    (notebooks/figures/performance/benchmark.py:63-67)."""
    import torch

    from cudams.similarity import CudaCosineGreedy
    from cudams.spectrum import Spectrum

    off = packed.offsets
    spectra = [Spectrum(packed.mz[off[i]:off[i + 1]], packed.intensities[off[i]:off[i + 1]])
               for i in range(packed.n_spectra)]
    cg = CudaCosineGreedy(tolerance=0.01, match_limit=2048)
    cg.matrix(spectra[:64], spectra[:64])
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = cg.matrix(spectra, spectra)
        times.append((time.perf_counter() - t0) * 1e3)
        del out
    best = min(times)
    n = packed.n_spectra
    log(f"matrix() wall {[round(t, 1) for t in times]} ms")
    return {"what": "CudaCosineGreedy(tolerance=0.01).matrix(spectra, spectra) -> np.recarray",
            "ms": best, "ms_runs": times, "pairs_per_s": n * n / (best / 1e3),
            "kernel_ms": cg.kernel_time * 1e3}


def bench_copy(a) -> dict:
    """A plain 1 GiB device-to-device copy: the calibration point for the
    FETCH_SIZE / WRITE_SIZE counters (their reading for a copy of known size)."""
    import torch

    n = 1 << 28  # float32 elements = 1 GiB
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(a.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    return {"metric": "device copy GB/s (counter calibration)", "value": 2 * 4 * n / (ms / 1e3) / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "copy_1GiB", "bytes_read": 4 * n,
                                            "bytes_written": 4 * n},
            "roofline": None, "cpu_baseline": None}


def bench_serving(a) -> dict:
    """Served library search: 500,000 library spectra (lognormal peaks, no
    clusters, seed 7) resident on the GPU and indexed once; requests of 1-256
    spectra (log-uniform) from a 4096-spectrum query pool (seed 8) arrive as a
    Poisson stream; each asks for the top-10 library hits per spectrum at
    tolerance 0.01.  Served = continuous batching (2 slots sharing a 6 GB
    output budget); plain = one search + top-k call per request, in arrival
    order, on the same index.  Latency = completion - scheduled arrival."""
    import numpy as np

    from cudams import synthetic
    from cudams.serving import LibraryIndex, pack_request, run_load
    from cudams.spectra import PackedSpectra

    n_lib = a.serve_library
    t0 = time.perf_counter()
    lib = synthetic.lognormal_spectra(n_lib, 7, cluster_fraction=0.0)
    log(f"serving: library {n_lib} spectra, {int(lib.offsets[-1])} peaks "
        f"({time.perf_counter() - t0:.1f} s)")
    ix = LibraryIndex(lib, tolerance=0.01)
    pool = synthetic.lognormal_spectra(4096, 8, cluster_fraction=0.0)

    def request(start, size):
        off = pool.offsets[start:start + size + 1]
        return pack_request(PackedSpectra(pool.mz[off[0]:off[-1]],
                                          pool.intensities[off[0]:off[-1]], off - off[0]))

    def stream(rate, seconds, seed):
        rng = np.random.default_rng(seed)
        n = max(1, int(rng.poisson(rate * seconds)))
        arrivals = np.sort(rng.uniform(0.0, seconds, n)) if rate > 0 else np.zeros(n)
        sizes = np.clip(np.floor(np.exp(rng.uniform(0.0, np.log(257.0), n))).astype(int), 1, 256)
        starts = rng.integers(0, pool.n_spectra - 256, n)
        return [request(s0, sz) for s0, sz in zip(starts, sizes)], arrivals

    # warm both paths (kernel attributes, allocator, threads)
    reqs, arr = stream(100.0, 0.5, 99)
    run_load("served", ix, reqs, arr)
    run_load("plain", ix, reqs, arr)
    rates = []
    # the workload's three rates, then a sustained rate between plain's and
    # served's saturation throughput (about 1,300 and 2,700 requests/s)
    for rate in (50.0, 200.0, 1000.0, 2000.0):
        # at least about 500 requests per rate, so that the p99 is not one of
        # the few largest requests alone
        seconds = max(a.serve_seconds, 500.0 / rate)
        reqs, arr = stream(rate, seconds, int(rate))
        row = {"rate_per_s": rate, "seconds": seconds}
        for mode in ("served", "plain"):
            r = run_load(mode, ix, reqs, arr)
            log(f"serving rate {rate}: {mode} p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms "
                f"max {r['max_ms']:.2f} ms setup {r['setup_ms']:.1f} ms "
                f"{r['requests_per_s']:.0f} req/s")
            row[mode] = {k: r[k] for k in ("p50_ms", "p99_ms", "p999_ms", "max_ms", "mean_ms",
                                           "setup_ms", "requests_per_s",
                                           "query_spectra_per_s", "requests")}
            if mode == "served":
                row[mode]["mean_batch_spectra"] = r["mean_batch_spectra"]
        rates.append(row)
    # saturation: a burst of requests all arriving at once; throughput =
    # requests / makespan
    reqs = [request(s0, sz) for s0, sz in zip(
        np.random.default_rng(6).integers(0, pool.n_spectra - 256, a.serve_burst),
        np.clip(np.floor(np.exp(np.random.default_rng(7).uniform(0.0, np.log(257.0),
                                                                 a.serve_burst))).astype(int),
                1, 256))]
    arr = np.zeros(len(reqs))
    sat = {}
    for mode in ("served", "plain"):
        r = run_load(mode, ix, reqs, arr)
        sat[mode] = {"requests_per_s": r["requests_per_s"],
                     "query_spectra_per_s": r["query_spectra_per_s"], "p99_ms": r["p99_ms"]}
        log(f"serving burst of {len(reqs)}: {mode} {r['requests_per_s']:.0f} req/s")
    sat["throughput_ratio"] = sat["served"]["requests_per_s"] / sat["plain"]["requests_per_s"]
    out = {"library_spectra": n_lib, "library_peaks": int(lib.offsets[-1]), "top_k": 10,
           "tolerance": 0.01, "request_spectra": "log-uniform 1-256", "rates": rates,
           "saturation": {"burst_requests": len(reqs), **sat}}
    if not a.no_cpu_baseline:
        from oracle import oracle

        cores = host_threads()
        csr = (lib.mz, lib.intensities, lib.offsets)
        cpu = {}
        for size in (1, 16):
            off = pool.offsets[:size + 1]
            q = (pool.mz[:off[-1]], pool.intensities[:off[-1]], off)
            t1 = time.perf_counter()
            oracle.cosine_greedy_matrix(q, csr, tolerance=0.01, nthreads=cores)
            cpu[str(size)] = (time.perf_counter() - t1) * 1e3
        out["cpu_request_latency_ms"] = {"cores": cores, "kind": "port", **cpu,
                                         "note": "oracle/cs_oracle.c float64 all pairs of the "
                                                 "request x library, top-k not included"}
    return out


def bench_batched(a) -> dict:
    """64 independent CosineGreedy matrices (seed 6): references and queries
    log-uniform in [16, 2048] x [16, 4096], lognormal peaks, tolerance 0.01 or
    0.1 per job, mz_power 0.5.  The autotuner picks the launch shape on the
    batch itself; then the graph-captured batch (one persistent lane-kernel
    launch for all jobs) is timed against one cs_cosine_greedy call per job,
    both with HIP events on the same stream."""
    import numpy as np
    import torch

    from cudams import synthetic
    from cudams.batch import BatchRunner, _DeviceJob, autotune, per_job, time_runner
    from cudams.device import Workspace

    jobs = synthetic.batched_jobs(seed=6)
    pairs = sum(j.refs.n_spectra * j.queries.n_spectra for j in jobs)
    log(f"batched: {len(jobs)} jobs, {pairs} pairs; autotuning")
    t0 = time.perf_counter()
    best, timings = autotune(jobs, replays=3)
    tune_s = time.perf_counter() - t0
    runner = BatchRunner(jobs, best)
    graph_ms = time_runner(runner, replays=max(3, a.steps))
    dev = runner.device
    ref = [_DeviceJob(j, dev) for j in jobs]
    ws = Workspace(dev)
    per_job(ref, ws)  # warm
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = max(3, a.steps)
    e0.record()
    for _ in range(reps):
        per_job(ref, ws)
    e1.record()
    e1.synchronize()
    per_ms = e0.elapsed_time(e1) / reps
    runner.run()
    same = all(all(np.array_equal(x, y) for x, y in zip(g, w))
               for g, w in zip(runner.results(), [d.result() for d in ref]))
    runner.close()
    log(f"batched: graph {graph_ms:.3f} ms, per-job launches {per_ms:.3f} ms, tuned {best}")
    return {"jobs": len(jobs), "pairs": pairs, "mz_power": 0.5, "tolerances": [0.01, 0.1],
            "graph_ms": graph_ms, "graph_pairs_per_s": pairs / (graph_ms / 1e3),
            "per_job_ms": per_ms, "per_job_pairs_per_s": pairs / (per_ms / 1e3),
            "speedup": per_ms / graph_ms, "identical_to_per_job": bool(same),
            "tuned": best.__dict__, "tuning_s": tune_s,
            "tuning_ms": {f"K{p.light_candidates}/rpb{p.refs_per_block}": round(v, 4)
                          for p, v in timings.items()}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--workload", choices=["cosine", "fingerprint", "copy", "batched"],
                    default="cosine")
    ap.add_argument("--light-candidates", type=int, default=0)
    ap.add_argument("--round-capacity", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the comparison of the whole matrix with the oracle")
    ap.add_argument("--no-matrix-wall", action="store_true",
                    help="skip the end-to-end timing of the drop-in matrix() call")
    ap.add_argument("--no-fingerprint", action="store_true",
                    help="skip the fingerprint section of the cosine line")
    ap.add_argument("--no-cpu-parallel", action="store_true",
                    help="skip the CPUParallelCosineGreedy timing of the cosine line "
                         "(--no-cpu-baseline skips it too)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-jobs section of the cosine line")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-search section of the cosine line")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-library", type=int, default=500_000)
    ap.add_argument("--serve-burst", type=int, default=1500)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); this build has no "
              "multi-GPU path to time.", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    res = {"cosine": bench_cosine, "fingerprint": bench_fingerprint, "copy": bench_copy,
           "batched": bench_batched_graph}[a.workload](a)
    if a.workload == "cosine" and not a.no_fingerprint:
        res["fingerprint"] = bench_fingerprint_measures(a)
    if a.workload == "cosine" and not a.no_batched:
        res["batched_jobs"] = bench_batched(a)
    if a.workload == "cosine" and not a.no_serving:
        res["serving"] = bench_serving(a)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
