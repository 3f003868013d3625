"""Build libfftm_mi.so in-tree with hipcc for gfx950.

    python -m fftm_amd._build          # compile every source (parallel), link
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfftm_mi.so")

SOURCES = ["mi_kernels_f32_a.hip", "mi_kernels_f32_b.hip", "mi_kernels_f32_c.hip", "mi_kernels_f32_d.hip",
           "mi_kernels_f64_a.hip", "mi_kernels_f64_b.hip", "mi_kernels_f64_c.hip", "mi_kernels_f64_d.hip",
           "mi_kernels_f32_s0.hip", "mi_kernels_f32_s1.hip", "mi_kernels_f32_s2.hip", "mi_kernels_f32_s3.hip",
           "mi_kernels_f64_s0.hip", "mi_kernels_f64_s1.hip", "mi_kernels_f64_s2.hip", "mi_kernels_f64_s3.hip",
           "mi_kernels_mixed_a.hip", "mi_kernels_mixed_b.hip", "mi_kernels_rs_512.hip", "mi_kernels_rs_1024.hip",
           "mi_kernels_mixed_s_a.hip", "mi_kernels_mixed_s_b.hip", "mi_kernels_mixed_s_c.hip",
           "mi_kernels_filter_f32.hip",
           "mi_kernels_filter_f64.hip",
           "mi_kernels_dispatch.hip", "mi_kernels_plane.hip", "mi_host.hip"]

# Packed FP32 VALU ops are disabled for every HIP translation unit (device side).
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_tests(verbose)
    return LIB


CPP_TESTS = [os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_ffts_cpp.cpp")]


def build_cpp_tests(verbose: bool = False):
    """The C++ drop-in test programs (host C++ over include/fftm_mi/*.hpp, linked to the library)."""
    for src in CPP_TESTS:
        out = os.path.splitext(src)[0] + ".bin"
        rel = os.path.relpath(HERE, os.path.dirname(src))
        cmd = [hipcc(), *HIPFLAGS, "-o", out, src, f"-L{HERE}", "-lfftm_mi", f"-Wl,-rpath,$ORIGIN/{rel}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {out}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
