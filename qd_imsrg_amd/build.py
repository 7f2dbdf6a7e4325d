"""Build libqdimsrg.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m qd_imsrg_amd.build   (also called by __graft_entry__.build()).
Every HIP translation unit is compiled with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libqdimsrg.so")
SOURCES = ["channels.cpp", "hf.cpp", "io.cpp", "kernels.hip", "boxes.hip", "gemm.hip", "capi.hip", "mlp.hip", "ode.hip", "observables.hip", "hf_device.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


HOST_OMP = {"hf.cpp"}   # host-only sources that use OpenMP


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    extra = ["-fopenmp"] if src in HOST_OMP else []
    cmd = [HIPCC, *FLAGS, *extra, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return out


FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
BIN = os.path.join(HERE, "bin")


def build_fortran(verbose: bool = False) -> str:
    """The iso_c_binding module and the run_magnus / gs_decouple drivers (amdflang)."""
    os.makedirs(BIN, exist_ok=True)
    mod = [FLANG, "-O2", "-c", os.path.join(FORTRAN, "qdimsrg_c.f90"), "-module-dir", OBJ,
           "-o", os.path.join(OBJ, "qdimsrg_c.o")]
    r = subprocess.run(mod, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed on qdimsrg_c.f90:\n{r.stderr}")
    for prog in ("run_magnus", "gs_decouple"):
        exe = os.path.join(BIN, prog)
        cmd = [FLANG, "-O2", os.path.join(FORTRAN, prog + ".f90"), "-I", OBJ, os.path.join(OBJ, "qdimsrg_c.o"),
               "-o", exe, "-L", HERE, "-lqdimsrg", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed on {prog}.f90:\n{r.stderr}")
    if verbose:
        print("built", exe)
    return exe


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
