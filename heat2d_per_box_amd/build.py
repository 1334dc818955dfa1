"""Build the HIP library ``libboxfgt_hip.so`` (gfx950) and the Fortran shim.

Compiles every source under ``csrc/`` with hipcc for gfx950 into
``_build/*.o`` and links ``libboxfgt_hip.so`` next to this file.  The
Fortran ``boxfgt`` drop-in (``fortran/boxfgt_hip.f90``) is compiled with
amdflang into ``libboxfgt_fortran.a`` when amdflang is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libboxfgt_hip.so")
FLIB = os.path.join(HERE, "libboxfgt_fortran.a")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep MFMA accumulators in ordinary VGPRs: the PW kernels feed one
    # product's accumulator to the next as an operand, and the AGPR form
    # copies every accumulator out and back on each loop iteration
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]

SOURCES = ["bfgt_tables.hip", "bfgt_pw.hip", "bfgt_near.hip", "bfgt_targets.hip", "bfgt_tree.hip", "bfgt_newtree.hip",
           "bfgt_plan.cpp"]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-soname,libboxfgt_hip.so", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    build_fortran_driver(verbose)
    return LIB


FDRIVER = os.path.join(BUILD, "test_boxfgt_dropin")


def build_fortran_driver(verbose: bool = False) -> str | None:
    """Link the Fortran known-answer driver (tests/fortran) against the drop-in."""
    src = os.path.join(os.path.dirname(HERE), "tests", "fortran", "test_boxfgt_dropin.f90")
    if not (os.path.exists(src) and os.path.exists(FLIB) and os.path.exists(FLANG)):
        return None
    cmd = [FLANG, "-O2", src, FLIB, f"-L{HERE}", "-lboxfgt_hip", "-Wl,-rpath,$ORIGIN/..", "-o", FDRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return FDRIVER


def build_fortran(verbose: bool = False) -> str | None:
    src = os.path.join(HERE, "fortran", "boxfgt_hip.f90")
    if not (os.path.exists(src) and os.path.exists(FLANG)):
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "boxfgt_hip_f.o")
    cmd = [FLANG, "-O2", "-fPIC", "-c", src, "-o", obj, "-J", BUILD]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    if os.path.exists(FLIB):
        os.remove(FLIB)
    subprocess.run(["ar", "rcs", FLIB, obj], check=True)
    return FLIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
