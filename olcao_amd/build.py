"""Build the gfx950 shared library olcao_amd/libolcao_gpu.so in-tree.

hipcc compiles each translation unit in parallel (the three-centre kernels
are split by bra angular momentum because their fully unrolled Obara-Saika
templates dominate compile time), then links against rocSOLVER/rocBLAS.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libolcao_gpu.so")
FLIB = os.path.join(HERE, "libolcao_fortran.so")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
ARCH = os.environ.get("OLCAO_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]

UNITS = [("kernels.hip", "kernels.o", []),
         ("kprim.hip", "kprim.o", []),
         ("k3c.hip", "k3c_la0.o", ["-DOLCAO_LA=0"]),
         ("k3c.hip", "k3c_la1.o", ["-DOLCAO_LA=1"]),
         ("k3c.hip", "k3c_la2.o", ["-DOLCAO_LA=2"]),
         ("k3c.hip", "k3c_la3.o", ["-DOLCAO_LA=3"]),
         ("ksolve.hip", "ksolve.o", []),
         ("ktrd.hip", "ktrd.o", []),
         ("ktrdb.hip", "ktrdb.o", []),
         ("kdc.hip", "kdc.o", []),
         ("kchol.hip", "kchol.o", []),
         ("kbloch.hip", "kbloch.o", []),
         ("noalloc.hip", "noalloc.o", []),
         ("runtime.cpp", "runtime.o", ["-x", "hip"])]


def _compile(unit):
    src, obj, extra = unit
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(SRC, src), "-o", os.path.join(BUILD, obj)]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return src, obj, time.time() - t


def build(verbose: bool = True, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 4)))
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, dt in ex.map(_compile, UNITS):
            if verbose:
                print(f"  compiled {src} -> {obj} in {dt:.1f}s", flush=True)
    objs = [os.path.join(BUILD, u[1]) for u in UNITS]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    build_fortran(verbose)
    build_h5(verbose)
    return LIB


def build_fortran(verbose: bool = True) -> str:
    """Fortran facade (ISO_C_BINDING) over the C-ABI, linked to the GPU
    library next to it."""
    src = os.path.join(HERE, "fortran", "olcao_gpu_facade.f90")
    tmp = FLIB + ".tmp"
    cmd = [FC, "-O2", "-fPIC", "-shared", "-module-dir", BUILD, src, "-o", tmp,
           f"-L{HERE}", "-lolcao_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, FLIB)
    if verbose:
        print(f"built {FLIB}", flush=True)
    build_fortran_driver(verbose)
    return FLIB


DRIVER_SRC = os.path.join(HERE, "..", "tests", "fortran", "facade_driver.f90")
DRIVER = os.path.join(HERE, "..", "tests", "fortran", "facade_driver")


def build_fortran_driver(verbose: bool = True):
    """The Fortran test program that uses the facade module at the
    reference driver's call sites (tests/fortran/facade_driver.f90), linked
    to the facade and GPU libraries."""
    if not os.path.exists(DRIVER_SRC):
        return None
    tmp = DRIVER + ".tmp"
    cmd = [FC, "-O2", "-module-dir", BUILD, "-I", BUILD, DRIVER_SRC, "-o", tmp, f"-L{HERE}",
           "-lolcao_fortran", "-lolcao_gpu", "-Wl,-rpath,$ORIGIN/../../olcao_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for the facade driver:\n{r.stderr[-4000:]}")
    os.replace(tmp, DRIVER)
    if verbose:
        print(f"built {DRIVER}", flush=True)
    return DRIVER


H5LIB = os.path.join(HERE, "libolcao_h5.so")
H5_PREFIX = "/opt/conda"


def build_h5(verbose: bool = True):
    """The setup's integral file in the reference's HDF5 layout
    (csrc/h5scf.cpp), host code over the HDF5 C library, linked to the GPU
    library next to it.  Skipped where no HDF5 C library is installed."""
    inc = os.path.join(H5_PREFIX, "include", "hdf5.h")
    if not os.path.exists(inc):
        if verbose:
            print(f"skipped {H5LIB}: no HDF5 C library under {H5_PREFIX}", flush=True)
        return None
    tmp = H5LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{H5_PREFIX}/include",
           os.path.join(SRC, "h5scf.cpp"), "-o", tmp, f"-L{HERE}", "-lolcao_gpu",
           f"-L{H5_PREFIX}/lib", "-lhdf5", "-Wl,--enable-new-dtags", "-Wl,-rpath,$ORIGIN",
           # the system C++ runtime before the HDF5 prefix's own, older one
           "-Wl,-rpath,/usr/lib/x86_64-linux-gnu", f"-Wl,-rpath,{H5_PREFIX}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5scf.cpp:\n{r.stderr[-4000:]}")
    os.replace(tmp, H5LIB)
    if verbose:
        print(f"built {H5LIB}", flush=True)
    return H5LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
