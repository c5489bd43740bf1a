"""Compile the native libraries of this package.

Usage: python cosine-similarity_amd/build.py [--jobs N]

* cudams/libcsamd.so: every .hip file under csrc/, compiled for gfx950 with
  hipcc (one object per file, in parallel) and linked into one shared library
  with a plain C ABI (include/cosine_similarity_amd.h).  hipcc cross-compiles,
  so this runs on a machine without a GPU.
* cudams/libcsamd_cpu.so: the host engine under csrc/cpu/ (g++, OpenMP), the
  compute of CPUParallelCosineGreedy; it needs no GPU.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "cudams", "libcsamd.so")
CPU_LIB = os.path.join(HERE, "cudams", "libcsamd_cpu.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: the float64 products and sums must round exactly as the
# matchms float64 reference writes them (no fused multiply-add).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU forms.
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


# host engine: float64 products and sums must round as matchms writes them
CPU_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wextra",
             "-Wno-unknown-pragmas", f"-I{INCLUDE}", f"-I{CSRC}"]


def cpu_sources():
    d = os.path.join(CSRC, "cpu")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".cpp"))


def build_cpu(out: str = CPU_LIB, extra=()) -> str:
    """g++ build of the host engine into ``out`` (``extra``: more flags)."""
    tmp = out + ".tmp"
    cmd = ["g++", *CPU_FLAGS, *extra, "-shared", "-o", tmp, *cpu_sources()]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build(jobs: int = 4, verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpu()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {CPU_LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
