"""Build libfdfd.so (HIP, gfx950) in-tree with hipcc.

Usage: python fdfd.jl_amd/build_lib.py [--jobs N]
Objects go to fdfd.jl_amd/_build/, the library to fdfd.jl_amd/lib/libfdfd.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfdfd.so")
ARCH = "gfx950"

SOURCES = ["k_operator.hip", "k_factor.hip", "k_bcr.hip", "k_mode.hip", "k_multi.hip",
           "runtime.cpp", "bcr.cpp", "multi.cpp",
           "eigen.cpp", "dense_eig.cpp", "batch.cpp", "mode.cpp", "modulated.cpp",
           "flux.cpp", "arnoldi.cpp", "vmode.cpp", "tune.cpp"]

# All device code is built without packed FP32 instructions.
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I" + os.path.join(HERE, "..", "include"),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: libfdfd needs the ROCm toolchain")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


# host-instrumented build for the sanitizer run (device code is not instrumented)
ASAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
ASAN_DIR = os.path.join(BUILD, "asan")
ASAN_EXE = os.path.join(LIBDIR, "asan_driver")


def _compile_asan(src: str) -> str:
    obj = os.path.join(ASAN_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, *ASAN_HOST, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_asan(jobs: int = 4, verbose: bool = False) -> str:
    """The library's sources with the host code instrumented by
    AddressSanitizer and UndefinedBehaviorSanitizer, linked into
    tests/asan_driver.cpp (run on a GPU by tests/test_gpu_asan.py)."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile_asan, SOURCES))
    driver = os.path.join(HERE, "..", "tests", "asan_driver.cpp")
    dobj = os.path.join(ASAN_DIR, "asan_driver.o")
    cmd = [hipcc(), "-O1", "-std=c++17", *ASAN_HOST, "-I" + os.path.join(HERE, "..", "include"),
           "-x", "c++", "-c", driver, "-o", dobj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    tmp = ASAN_EXE + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", *ASAN_HOST, dobj, *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ASAN_EXE)
    if verbose:
        print("built", ASAN_EXE)
    return ASAN_EXE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
