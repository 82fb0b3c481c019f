"""Build recipe of the gfx950 library (librpfvm.so, in-tree).

    python -m overheatingproblem_amd.build

Every HIP translation unit is compiled for gfx950 only, with FMA contraction off
(the FP64 form must round like the reference's x86-64 build) and packed FP32 ops
disabled; the objects are linked into overheatingproblem_amd/lib/librpfvm.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.environ.get("RPFVM_LIB_DIR", os.path.join(PKG, "lib"))
LIB = os.path.join(LIBDIR, "librpfvm.so")
PROBE = os.path.join(LIBDIR, "hbm_probe")
DRIVER = os.path.join(LIBDIR, "rp_fvm_gpu")
BUILD = os.environ.get("RPFVM_BUILD_DIR", os.path.join(PKG, "_build"))

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
          "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable"]

# extra compiler flags, e.g. RPFVM_HIPFLAGS="-DRPFVM_FLUX_WAVES=4"
EXTRA = os.environ.get("RPFVM_HIPFLAGS", "").split()

SOURCES = ["fvm_kernels.hip", "fvm_mixed_kernels.hip", "fvm_batch_kernels.hip", "fvm_api.cpp",
           "fvm_batch_api.cpp", "fvm_noalloc.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *EXTRA, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the copy kernel that calibrates rocprofv3's HBM byte counters
    cmd = [HIPCC, *COMMON, "-xhip", os.path.join(CSRC, "hbm_probe.cpp"), "-o", PROBE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # the C driver (host code only, plain gcc; finds librpfvm.so next to itself)
    cmd = ["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-Wall", "-I",
           os.path.join(ROOT, "include"), os.path.join(PKG, "driver", "rp_fvm_gpu.c"),
           "-L", LIBDIR, "-lrpfvm", "-Wl,-rpath,$ORIGIN", "-lm", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}, {PROBE} and {DRIVER}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
