"""Shared test setup.

* marker ``gpu``: needs a ROCm GPU (run on the MI355X); everything else runs
  on the CPU.
* puts the repository root (for ``oracle``) and ``cosine-similarity_amd``
  (for the ``cudams`` package) on sys.path.
"""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cosine-similarity_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")
NATIVE = os.path.join(REPO, "tests", "native")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    o.lib()
    return o


def build_tile_harness() -> str:
    so = os.path.join(NATIVE, "libtile_harness.so")
    src = os.path.join(NATIVE, "tile_harness.cpp")
    deps = [src] + [os.path.join(PKG, "csrc", h) for h in ("cg_lane.h", "cg_tile.h")]
    if not os.path.exists(so) or any(os.path.getmtime(so) < os.path.getmtime(d) for d in deps):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        f"-I{os.path.join(PKG, 'csrc')}", "-o", so, src], check=True)
    return so


@pytest.fixture(scope="session")
def tile_harness():
    import ctypes

    L = ctypes.CDLL(build_tile_harness())
    d, p, i = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    L.tile_scan.argtypes = [i, p, p, p, i, p, p, p, i, d, d, d, d, d, p, p, p, p]
    L.coarse_tolerance.argtypes = [d, d, d]
    L.coarse_tolerance.restype = ctypes.c_float
    return L


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a ROCm GPU")
    from cudams import _native

    _native.load()
    return torch.device("cuda:0")
