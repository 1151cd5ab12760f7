"""Build libmfe.so (gfx950) and the Fortran interface objects in-tree.

    python -m molsim_amd.build            # or molsim_amd.build.build()

The device code is compiled for gfx950 only.  The host-side table builder
(mfe_tables.cpp) is compiled by g++ with -ffp-contract=off so its FP64
arithmetic rounds every product separately, like the reference's x86-64
build; the HIP sources keep hipcc's default contraction except where a
kernel disables it locally.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FORTRAN = os.path.join(PKG, "fortran")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmfe.so")

HIP_SOURCES = ["mfe_engine.hip", "mfe_kernels.hip", "mfe_fft.hip", "mfe_batch.hip",
               "mfe_fm.hip"]
CXX_SOURCES = ["mfe_tables.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay disabled (v_pk_* FP32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" +
                           r.stdout[-4000:] + r.stderr[-8000:])
    return r


def _compile_one(args):
    src, obj, kind = args
    if kind == "hip":
        cmd = [HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj]
    else:
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off",
               "-fno-fast-math", "-I", INCLUDE, "-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source (always: this is the does-it-build check)."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append((os.path.join(CSRC, s), os.path.join(BUILD, s + ".o"), "hip"))
    for s in CXX_SOURCES:
        jobs.append((os.path.join(CSRC, s), os.path.join(BUILD, s + ".o"), "cxx"))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        objs = list(ex.map(_compile_one, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_fortran()
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran() -> None:
    """Compile the ISO_C_BINDING module and the example host program."""
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        return
    out = os.path.join(BUILD, "fortran")
    os.makedirs(out, exist_ok=True)
    mod = os.path.join(FORTRAN, "mfe_mod.F90")
    host = os.path.join(FORTRAN, "mfe_host.F90")
    if not os.path.exists(mod):
        return
    _run([fc, "-O2", "-c", mod, "-o", os.path.join(out, "mfe_mod.o")], cwd=out)
    if os.path.exists(host):
        _run([fc, "-O2", "-c", host, "-o", os.path.join(out, "mfe_host.o")], cwd=out)
        _run([fc, "-O2", "-o", os.path.join(PKG, "mfe_host"),
              os.path.join(out, "mfe_host.o"), os.path.join(out, "mfe_mod.o"),
              "-L" + PKG, "-lmfe", "-Wl,-rpath,$ORIGIN"], cwd=out)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
