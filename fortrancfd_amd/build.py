"""Build libfcfd_hip.so (gfx950) in-tree with hipcc.

    python -m fortrancfd_amd.build [--verbose]

Every HIP translation unit is compiled with -ffp-contract=off (the fluid path reproduces the
reference's rounding) and with the packed-FP32 feature disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(OUT_DIR, "libfcfd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["api.hip", "fluid.hip", "chem.hip", "chem_lane.hip", "tran.hip", "aux.hip", "mlp.hip", "surrogate.hip", "mech_io.cpp", "mech_text.cpp", "kin_build.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          f"-I{os.path.join(REPO, 'include')}", f"-I{CSRC}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *COMMON]
        if src.endswith(".hip"):
            cmd += ["-x", "hip", *HIP_FLAGS]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        for r in list(pool.map(lambda c: subprocess.run(c), cmds)):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_fortran(verbose)
    build_h5(verbose)
    return LIB


HDF5_PREFIX = os.environ.get("FCFD_HDF5", "/opt/conda")
LIB_H5 = os.path.join(OUT_DIR, "libfcfd_h5.so")


def build_h5(verbose: bool = False) -> str | None:
    """The snapshot writer/reader (csrc/h5io.cpp, host only) over the HDF5 C library of
    $FCFD_HDF5 (default /opt/conda) -> _lib/libfcfd_h5.so; skipped when hdf5.h is absent."""
    inc, libdir = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"hdf5.h not found under {inc}: libfcfd_h5.so not built", flush=True)
        return None
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{inc}",
           os.path.join(CSRC, "h5io.cpp"), "-o", LIB_H5 + ".tmp", f"-L{libdir}", "-lhdf5",
           f"-Wl,-rpath,{libdir}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB_H5 + ".tmp", LIB_H5)
    return LIB_H5


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = False) -> str:
    """The Fortran binding module (fortran/fcfd.f90) -> _lib/fortran/{fcfd.mod, fcfd.o}."""
    out = os.path.join(OUT_DIR, "fortran")
    os.makedirs(out, exist_ok=True)
    cmd = [FLANG, "-c", "-O2", os.path.join(PKG, "fortran", "fcfd.f90"), "-module-dir", out,
           "-o", os.path.join(out, "fcfd.o")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
