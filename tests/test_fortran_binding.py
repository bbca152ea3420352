"""The Fortran binding (orctm_amd/fortran/orctm_nh_c.f90) compiled with LLVM flang, on the CPU.

* It compiles as ORCTM compiles (preprocessed, -r8), with and without -DCG2D4NONHY, against test
  stand-ins of the model's modules declared as ORCTM declares them (DZW a plain array, every field a
  REAL POINTER), and links against liborctm_nh.so.
* The Fortran host runs here, with no GPU: it reads the model's arrays, and petsc_start reports the C
  ABI's status (no device) and stops, as the binding's error path says.
The GPU run of the same host is tests/test_gpu_fortran.py.
"""
import importlib.util
import os
import subprocess

import pytest

from fortran_host import FBUILD, inputs, write_input
from orctm_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


def _fbuild():
    spec = importlib.util.spec_from_file_location("fortran_build", os.path.join(HERE, "fortran", "build.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_binding_compiles_and_links(tmp_path):
    fb = _fbuild()
    if fb.flang() is None:
        pytest.skip("no Fortran compiler")
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("liborctm_nh.so is not built")
    built = fb.build(out=str(tmp_path))
    assert set(built) == {"plain", "cg2d", "restart"}
    for exe in built.values():
        assert os.access(exe, os.X_OK)


def test_host_reports_no_device(tmp_path):
    exe = os.path.join(FBUILD, "plain", "orctm_nh_fdriver")
    if not os.path.exists(exe):
        pytest.skip("the Fortran host is not built")
    lib = abi.load()
    import ctypes

    n = ctypes.c_int32(0)
    lib.orctm_nh_device_count(ctypes.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present; tests/test_gpu_fortran.py runs the host there")
    fin = str(tmp_path / "in.bin")
    write_input(fin, *inputs())
    r = subprocess.run([exe, fin, str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "orctm_nh_create failed, status" in r.stdout, r.stdout
    assert int(r.stdout.split("status")[-1].split()[0]) == abi.ERR_NODEVICE
