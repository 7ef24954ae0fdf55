"""Builds dcsr_amd/libdcsr.so (hipcc, gfx950) and the CPU checkers in oracle/.

Usage:  python -m dcsr_amd.build   (or __graft_entry__.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdcsr.so")

HIP_SOURCES = ["prims.hip", "gsort.hip", "mtx.hip", "csrcmp.hip", "matrix.hip", "update.hip", "spmv.hip", "spgemm.hip", "hyb.hip", "hybstore.hip", "batch.hip",
               "server.hip", "noalloc.hip"]
CXX_SOURCES: list[str] = []

ARCH = os.environ.get("DCSR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # exact FP32 rounding of every add and multiply (no contraction to FMA)
    "-ffp-contract=off",
    # no packed FP32 math in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + CXX_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "shim", "ref_style_driver.cpp")
DRIVER_BIN = os.path.join(ROOT, "tests", "shim", "ref_style_driver")


def driver_cmd(out: str) -> list[str]:
    """hipcc line that builds the reference-style driver (test infrastructure:
    a caller of the C++ drop-in shim include/dcsr_matrix.hpp) against
    libdcsr.so, found next to the tree at run time."""
    return [HIPCC, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
            DRIVER_SRC, "-o", out, "-L", HERE, "-ldcsr", "-Wl,-rpath,$ORIGIN/../../dcsr_amd"]


def build_driver(verbose: bool = False) -> None:
    r = subprocess.run(driver_cmd(DRIVER_BIN + ".tmp"), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER_BIN + ".tmp", DRIVER_BIN)
    if verbose:
        print("built", DRIVER_BIN)


def build_oracle(verbose: bool = False) -> None:
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(r.stdout.strip())


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_driver(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
