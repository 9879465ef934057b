"""Build recipe for libo3d.so (HIP, gfx950) and the Fortran drop-in modules.

Every source is compiled for gfx950 only.  FP contraction is off so that
the FP64 stencils reproduce the reference's rounding (amdflang on x86-64
emits no FMA); packed FP32 instructions are disabled.
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
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libo3d.so")
FLIB = os.path.join(PKG, "libo3d_fortran.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("O3D_FC", "/opt/rocm/bin/amdflang")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]

SOURCES = ["o3d_stencil.hip", "o3d_sor.hip", "o3d_mg.hip", "o3d_nn.hip", "o3d_api.hip"]
# Modules that compile on their own.  differential_operators, les_turbulence
# and integration use the reference's initialization module (as the
# reference's own versions do); they are built together with it by
# `make -C oracle dropin`.
FORTRAN_SOURCES = ["o3d_capi.f90", "derivation.f90", "poisson.f90"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING drop-in modules (same module and procedure names as
    the reference), linked against libo3d.so."""
    srcs = [os.path.join(FSRC, s) for s in FORTRAN_SOURCES if os.path.exists(os.path.join(FSRC, s))]
    if not srcs or not os.path.exists(FC):
        return ""
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for s in srcs:
        obj = os.path.join(mod, os.path.basename(s).replace(".f90", ".o"))
        _run([FC, "-O2", "-fPIC", "-module-dir", mod, "-c", s, "-o", obj], verbose)
        objs.append(obj)
    tmp = FLIB + ".tmp"
    _run([FC, "-shared", "-fPIC", *objs, "-L" + PKG, "-lo3d", "-Wl,-rpath,$ORIGIN", "-o", tmp],
         verbose)
    os.replace(tmp, FLIB)
    return FLIB


def build_oracle(verbose: bool = False, reference: str = "/root/reference") -> None:
    """The oracle: the C restatement always, the reference modules (and the
    Fortran drop-in built against them) when the reference sources are
    present at build time."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "oracle"], verbose)
    if os.path.isdir(os.path.join(reference, "src")) and shutil.which("make"):
        _run(["make", "-s", "-C", odir, "ref", f"REF={reference}"], verbose)
        if os.path.exists(LIB):
            _run(["make", "-s", "-C", odir, "dropin", f"REF={reference}"], verbose)


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_fortran(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
