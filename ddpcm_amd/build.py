"""Builds libddx.so (HIP, gfx950) and the Fortran drop-in driver in-tree.

    python -m ddpcm_amd.build            # library + Fortran driver
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
FSRC = os.path.join(PKG, "fortran")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libddx.so")
DRIVER = os.path.join(PKG, "ddx_main")
TEST_PROG = os.path.join(ROOT, "tests", "fortran", "jd_module")
PKG_RPATH = "$ORIGIN/../../ddpcm_amd"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLANG = "/opt/rocm/lib/llvm/bin/flang"
ARCH = os.environ.get("DDX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I" + os.path.join(ROOT, "include"),
]

SOURCES = ["ddx_api.hip", "k_scan.hip", "k_setup.hip", "k_gemm.hip", "k_sweep.hip",
           "k_solver.hip", "k_coulomb.hip", "k_forces.hip", "k_guess.hip", "k_ops.hip", "k_fmm.hip",
           "tables.cpp"]


# per-source additions: the matrix products keep MFMA accumulators in VGPRs
# (the AGPR form copies every loop-carried accumulator in and out per
# iteration and doubles the register count)
SOURCE_FLAGS = {"k_gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-x", "c++", "-I" + os.path.join(ROOT, "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """The Fortran module (reference-named procedures over the C-ABI) and the
    drop-in driver reading Input.txt, linked against libddx.so."""
    if not os.path.exists(FLANG):
        return None
    srcs = [os.path.join(FSRC, f) for f in ("ddx_cabi.f90", "ddcosmo_gpu.f90", "ddx_main.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(OBJ, exist_ok=True)
    cmd = [FLANG, "-O2", "-module-dir", OBJ] + srcs + [
        "-o", DRIVER, "-L" + PKG, "-lddx", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", DRIVER)
    # the module's test program (tests/fortran), linked the same way
    tsrc = os.path.join(ROOT, "tests", "fortran", "jd_module.f90")
    if os.path.exists(tsrc):
        tmod = os.path.join(OBJ, "tests")
        os.makedirs(tmod, exist_ok=True)
        cmd = [FLANG, "-O2", "-module-dir", tmod] + srcs[:2] + [tsrc, "-o", TEST_PROG,
                                                                "-L" + PKG, "-lddx",
                                                                "-Wl,-rpath," + PKG_RPATH]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n%s\n%s" % (r.stdout, r.stderr))
        if verbose:
            print("built", TEST_PROG)
    return DRIVER


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
