"""Builds the Fortran host of the binding (orctm_amd/fortran/orctm_nh_c.f90) with LLVM flang:

    python tests/fortran/build.py

Three executables, for the grid the GPU tests feed them (the shared "regional" geometry: KE = 13, not
periodic), compiled as ORCTM compiles (-r8 -> -fdefault-real-8, preprocessed):
    _build/plain/orctm_nh_fdriver     the shipped configuration
    _build/cg2d/orctm_nh_fdriver      with -DCG2D4NONHY (the 2-D bottom-pressure solve)
    _build/restart/orctm_nh_frestart  the resident pressure saved and restored with the restart files
All link liborctm_nh.so from orctm_amd/ through an $ORIGIN-relative run path, so the tree can move.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BINDING = os.path.join(REPO, "orctm_amd", "fortran", "orctm_nh_c.f90")
LIBDIR = os.path.join(REPO, "orctm_amd")
OUT = os.path.join(HERE, "_build")
FLANG_CANDIDATES = ("/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang")

# the grid the executables are compiled for (compile-time KE and periodicity, as in MO_PARAM1)
TEST_KE, TEST_CYCLIC_X, TEST_CYCLIC_Y = 13, 0, 0
# variant: (extra flags, driver source, executable)
VARIANTS = {"plain": ([], "driver.F90", "orctm_nh_fdriver"),
            "cg2d": (["-DCG2D4NONHY"], "driver.F90", "orctm_nh_fdriver"),
            "restart": ([], "restart_driver.F90", "orctm_nh_frestart")}


def flang() -> str | None:
    for c in FLANG_CANDIDATES:
        if os.path.exists(c):
            return c
    return shutil.which("flang")


def exe_path(variant: str) -> str:
    return os.path.join(OUT, variant, VARIANTS[variant][2])


def build(out: str = OUT, verbose: bool = False) -> dict:
    fc = flang()
    if fc is None:
        raise RuntimeError("no Fortran compiler (amdflang / flang) found")
    flags = ["-cpp", "-O2", "-fdefault-real-8", f"-DTEST_KE={TEST_KE}", f"-DTEST_CYCLIC_X={TEST_CYCLIC_X}",
             f"-DTEST_CYCLIC_Y={TEST_CYCLIC_Y}"]
    built = {}
    for variant, (defs, driver, name) in VARIANTS.items():
        d = os.path.join(out, variant)
        os.makedirs(d, exist_ok=True)
        objs = []
        for src in (os.path.join(HERE, "host_modules.F90"), BINDING, os.path.join(HERE, driver)):
            obj = os.path.join(d, os.path.splitext(os.path.basename(src))[0] + ".o")
            cmd = [fc, *flags, *defs, "-module-dir", d, "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, cwd=d)
            objs.append(obj)
        exe = os.path.join(d, name)
        rel = os.path.relpath(LIBDIR, d)
        cmd = [fc, *objs, "-o", exe, f"-L{LIBDIR}", "-lorctm_nh", f"-Wl,-rpath,$ORIGIN/{rel}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=d)
        built[variant] = exe
    return built


if __name__ == "__main__":
    for k, v in build(verbose="-v" in sys.argv).items():
        print(k, v)
