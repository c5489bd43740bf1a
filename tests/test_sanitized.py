"""The host code built a second time under AddressSanitizer and
UndefinedBehaviorSanitizer (gcc, -fno-sanitize-recover=all): the host engine
of CPUParallelCosineGreedy (csrc/cpu/, with cg_tile.h / cg_lane.h /
cg_check.h), the lane/tile harness (tests/native/tile_harness.cpp, the same
headers the GPU kernels include) and the oracle (oracle/cs_oracle.c).  A
child Python process loads the instrumented libraries with the ASan runtime
preloaded, runs them on ragged, clustered, tied, shifted and empty inputs
and compares them with each other; any out-of-bounds access, overflow or
other undefined behaviour aborts it."""
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import NATIVE, PKG, REPO

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.fixture(scope="module")
def sanitized_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    csrc = os.path.join(PKG, "csrc")
    oracle_so, harness_so, cpu_so = (str(d / n) for n in ("liboracle.so", "libharness.so",
                                                         "libcpu.so"))
    _run(["gcc", *SAN, "-fPIC", "-fopenmp", "-std=c11", "-ffp-contract=off", "-shared", "-o",
          oracle_so, os.path.join(REPO, "oracle", "cs_oracle.c"), "-lm"])
    _run(["g++", *SAN, "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-I{csrc}", "-o",
          harness_so, os.path.join(NATIVE, "tile_harness.cpp")])
    sys.path.insert(0, PKG)
    import build as pkg_build  # cosine-similarity_amd/build.py

    pkg_build.build_cpu(cpu_so, extra=SAN)
    return oracle_so, harness_so, cpu_so


SCRIPT = textwrap.dedent("""
    import ctypes, sys
    import numpy as np
    repo, pkg, tests, oracle_so, harness_so, cpu_so = sys.argv[1:7]
    sys.path[:0] = [repo, pkg, tests]
    from oracle import oracle
    oracle._LIB_PATH = oracle_so
    from cudams import _native, synthetic
    from cudams.spectra import PackedSpectra
    _native.load_cpu(cpu_so)
    from cudams.similarity import CPUParallelCosineGreedy
    from cudams.spectrum import Spectrum
    import test_tile_logic as ttl

    L = ctypes.CDLL(harness_so)
    d, p, i = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.tile_scan.argtypes = [i, p, p, p, i, p, p, p, i, d, d, d, d, d, p, p, p, p]

    def spectra(x):
        o = x.offsets
        return [Spectrum(x.mz[o[k]:o[k + 1]], x.intensities[o[k]:o[k + 1]])
                for k in range(x.n_spectra)]

    def engine_vs_oracle(refs, ques, **kw):
        got = CPUParallelCosineGreedy(**kw).matrix(spectra(refs), spectra(ques))
        okw = {"tolerance": kw.get("tolerance", 0.1), "shift": kw.get("shift", 0.0),
               "mz_power": kw.get("mz_power", 0.0),
               "intensity_power": kw.get("intensity_power", 1.0)}
        want = oracle.cosine_greedy_matrix((refs.mz, refs.intensities, refs.offsets),
                                           (ques.mz, ques.intensities, ques.offsets), **okw)
        assert np.array_equal(got["matches"], want["matches"])
        assert np.nanmax(np.abs(got["score"] - want["score"])) <= 1e-5

    rng = np.random.default_rng(3)
    clustered = synthetic.lognormal_spectra(96, 5)
    engine_vs_oracle(clustered, clustered, tolerance=0.01)
    for K in (1, 8, 16):
        ttl.check(L, oracle, K, clustered, clustered, 0.01)
    grid = ttl.grid_spectra(rng, 70)
    for tol in (0.25, 0.5, 0.0):
        engine_vs_oracle(grid, grid, tolerance=tol)
        ttl.check(L, oracle, 8, grid, grid, tol)
    mzs = [np.sort(rng.uniform(100, 900, rng.integers(0, 30))) for _ in range(40)]
    f64 = PackedSpectra.from_arrays(mzs, [rng.random(len(m)) + 0.01 for m in mzs])
    engine_vs_oracle(f64, f64, tolerance=0.05, shift=1.5, mz_power=0.5, intensity_power=2.0)
    ttl.check(L, oracle, 4, f64, f64, 0.05, 1.5, 0.5, 2.0)
    oracle.fingerprint_matrix(rng.integers(0, 2, (17, 70)), rng.integers(0, 2, (9, 70)), "cosine")
    print("sanitized ok")
""")


def test_host_code_under_asan_and_ubsan(sanitized_libs):
    oracle_so, harness_so, cpu_so = sanitized_libs
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-c", SCRIPT, REPO, PKG, os.path.join(REPO, "tests"),
                        oracle_so, harness_so, cpu_so], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and "sanitized ok" in r.stdout, r.stdout[-4000:] + r.stderr[-4000:]
