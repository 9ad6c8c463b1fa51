"""Build the HIP library (gfx950) in-tree: rsdft_amd/librsdft_gpu.so.

Every .hip file under rsdft_amd/csrc is compiled with hipcc to an object in
rsdft_amd/_build/, then linked with rocBLAS and rocSOLVER into one shared
library that exports exactly the C-ABI of include/rsdft_gpu.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librsdft_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(BUILD, "fortran")
DRIVER = os.path.join(FBUILD, "sweep_driver")


def build_fortran(verbose: bool = False) -> str:
    """Compile the ISO_C_BINDING module and the sweep driver against the library."""
    os.makedirs(FBUILD, exist_ok=True)
    mod_obj = os.path.join(FBUILD, "rsdft_gpu_mod.o")
    steps = [
        [FLANG, "-O2", "-fopenmp", "-c", os.path.join(FDIR, "rsdft_gpu_mod.f90"), "-J", FBUILD,
         "-o", mod_obj],
        [FLANG, "-O2", "-fopenmp", os.path.join(FDIR, "sweep_driver.f90"), "-J", FBUILD,
         "-I", FBUILD, mod_obj, "-o", DRIVER, "-L" + HERE, "-lrsdft_gpu",
         "-Wl,-rpath," + HERE, "-Wl,-rpath,/opt/rocm/lib"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
