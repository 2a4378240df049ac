"""Build recipe of the native libraries of exoplasim_amd (gfx950 only).

``build_library()`` compiles the HIP sources in csrc/ into
``exoplasim_amd/libexosht.so`` with hipcc.  It runs on a machine without a
GPU (hipcc cross-compiles for gfx950).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libexosht.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["sht_tables.hip", "sht_transforms.hip", "sht_transforms_m0.hip",
           "sht_transforms_m1.hip", "sht_transforms_m2.hip", "sht_transforms_m3.hip",
           "sht_direct_terms.hip", "sht_capi.hip", "sht_ensemble.hip", "si_solve.hip",
           "dyn_calcgp.hip", "sht_multiproc.hip", "sht_output.hip", "rad_lwr.hip",
           "rad_mlp.hip"]
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every source to an object (in parallel), then link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJDIR, exist_ok=True)
    flags = [f for f in HIPFLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *flags, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        return obj, subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    failed = [r for _, r in results if r.returncode != 0]
    if failed:
        for r in failed:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libexosht.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPFLAGS, "-o", tmp, *[o for o, _ in results]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libexosht.so")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BIN = os.path.join(FORTRAN, "bin")
SHIM = ["exo_legmod_capi.f90", "exo_transforms.f90"]


def build_fortran_driver(driver: str, verbose: bool = False) -> list:
    """Compile the Fortran drop-in (exo_legmod_capi.f90 + exo_transforms.f90)
    with a test program into fortran/bin/driver_r8 (-fdefault-real-8, the
    MOST_PREC=8 model) and driver_r4, linked to libexosht.so."""
    fc = os.environ.get("EXO_FC") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    os.makedirs(FORTRAN_BIN, exist_ok=True)
    outs = []
    for tag, flags in (("r8", ["-fdefault-real-8"]), ("r4", [])):
        mod = os.path.join(FORTRAN_BIN, f"mod_{tag}")
        os.makedirs(mod, exist_ok=True)
        exe = os.path.join(FORTRAN_BIN, f"driver_{tag}")
        cmd = [fc, "-O2", *flags, f"-J{mod}", *[os.path.join(FORTRAN, s) for s in SHIM],
               driver, "-o", exe, f"-L{HERE}", "-lexosht", f"-Wl,-rpath,{HERE}",
               "-Wl,-rpath,$ORIGIN/../.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("amdflang failed building the Fortran drop-in driver")
        outs.append(exe)
    return outs


if __name__ == "__main__":
    print(build_library(verbose=True))
