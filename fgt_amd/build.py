"""Builds ``fgt_amd/libfgt_amd.so`` in-tree with hipcc for gfx950.

Usage: ``python -m fgt_amd.build`` (also called by ``__graft_entry__.build``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfgt_amd.so")

SOURCES = ["plan.cpp", "host_api.cpp", "tree.cpp", "retree.cpp", "api.cpp", "kernels.hip", "kernels_pw3.hip", "kernels_direct3.hip",
           "kernels_targets.hip", "noalloc.hip"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-fopenmp",
    # a variable read before it is set (e.g. used in its own initialiser)
    # fails the build
    "-Werror=uninitialized", "-Werror=sometimes-uninitialized",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        cmd = [hipcc(), *HIP_FLAGS, "-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    build_fortran_driver(verbose)
    return OUT


FORTRAN_DRIVER = os.path.join(HERE, "_build", "drop_in_driver")


def flang():
    for cand in (shutil.which("amdflang"), "/opt/rocm/bin/amdflang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran_driver(verbose: bool = False):
    """The Fortran drop-in test driver (tests/fortran), linked against the
    library through the reference's implicit ``boxfgt`` interface and the
    ``fgt_amd`` module (include/fgt_amd.f90).  Skipped without a Fortran
    compiler (the tests that run it skip too)."""
    fc = flang()
    if fc is None:
        return None
    root = os.path.dirname(HERE)
    moddir = os.path.join(HERE, "_build", "fmod")
    os.makedirs(moddir, exist_ok=True)
    cmd = [fc, "-O2", "-J", moddir, os.path.join(root, "include", "fgt_amd.f90"),
           os.path.join(root, "tests", "fortran", "drop_in_driver.f90"), "-o", FORTRAN_DRIVER,
           "-L", HERE, "-lfgt_amd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=moddir)
    return FORTRAN_DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
