"""Build libifdda_hip.so (gfx950) and the Fortran drop-in shim, in-tree.

    python if-dda_amd/ifdda/build.py

hipcc compiles every HIP source for gfx950 with packed-FP32 instructions
disabled; amdflang compiles the iso_c_binding module against it.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libifdda_hip.so")
FLIB = os.path.join(PKG, "libifdda_dropin.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=on",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_hip():
    """One object per source, compiled in parallel (each source is its own
    translation unit either way), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def obj(src):
        o = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        _run([hipcc, *cflags, "-c", "-o", o, src])
        return o

    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(obj, srcs))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    src = os.path.join(PKG, "fortran", "ifdda_dropin.f90")
    moddir = os.path.join(PKG, "fortran", "_mod")
    os.makedirs(moddir, exist_ok=True)
    tmp = FLIB + ".tmp"
    _run([fc, "-O2", "-fPIC", "-shared", f"-J{moddir}", "-o", tmp, src,
          f"-L{PKG}", "-lifdda_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    return FLIB


def build_all():
    build_hip()
    if os.path.exists(os.path.join(PKG, "fortran", "ifdda_dropin.f90")):
        build_fortran()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
