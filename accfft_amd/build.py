"""Builds the gfx950 shared library ``accfft_amd/libaccfft_amd.so`` in tree.

Every HIP translation unit under ``accfft_amd/csrc`` is compiled with hipcc
for gfx950 (in parallel) and linked into one C-ABI library whose symbols are
declared in ``include/accfft_amd.h``.

Usage: ``python -m accfft_amd.build`` (or ``build()`` from Python).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaccfft_amd.so")

ARCH = os.environ.get("ACCFFT_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-function",
    "-Wno-pass-failed",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
if os.environ.get("ACCFFT_AMD_SWEEP", "0") not in ("", "0"):
    # measurement build: tile / order overrides read from the environment
    FLAGS.append("-DACCFFT_AMD_SWEEP")
SOURCES = ["fft_f32.hip", "fft_f64.hip", "fft_split.hip", "transpose.hip", "capi.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


ROOT = os.path.dirname(HERE)
DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_steps.cpp")
DROPIN_BIN = os.path.join(ROOT, "tests", "cpp", "dropin_steps")


def build_dropin_example() -> str:
    """Builds tests/cpp/dropin_steps, a program written against the
    reference's C++ API, from the drop-in headers in include/."""
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           DROPIN_SRC, "-L", HERE, "-laccfft_amd", "-Wl,-rpath,$ORIGIN/../../accfft_amd",
           "-o", DROPIN_BIN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in example failed to build:\n" + r.stderr)
    os.replace(DROPIN_BIN + ".tmp", DROPIN_BIN)
    return DROPIN_BIN


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "4")) or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_dropin_example()
    if verbose:
        print(f"built {LIB} and {DROPIN_BIN}")
    return LIB


if __name__ == "__main__":
    build()
