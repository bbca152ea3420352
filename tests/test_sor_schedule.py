"""The device's temporally blocked red-black SOR schedule (orctm_amd/csrc/nh_sor_tb.h: several half-sweeps
per launch in LDS, the cyclic halo read through a periodic index map, Jacobi order within a half-sweep,
ping-pong buffers, one exchange at the end) stepped thread by thread on the host, against the oracle's
sequential TRONEU loop (troneu.f90:104-131, oracle/nh_oracle_sor.c).

Bar: bit-exact Z1O for every compiled depth, for the regular and the seam kernels, on open, closed and
periodic domains, odd cyclic periods, domains smaller than one launch's halo, and 0 or 1 sweeps. The
GPU tests run the same kernels on the device (tests/test_gpu_sor.py); this one runs everywhere hipcc does.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return None


def test_blocked_schedule_matches_sequential_sweeps(tmp_path):
    hipcc, gcc = _hipcc(), shutil.which("gcc")
    if hipcc is None or gcc is None:
        pytest.skip("needs hipcc and gcc")
    objs = []
    for src in ("nh_oracle_ocv.c", "nh_oracle_sor.c"):
        o = tmp_path / (src + ".o")
        subprocess.run([gcc, "-O1", "-std=c11", "-ffp-contract=off", "-c", "-o", str(o),
                        os.path.join(ROOT, "oracle", src)], check=True)
        objs.append(str(o))
    h = tmp_path / "harness.o"
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O1", "-std=c++17", "-ffp-contract=off", "-w", "-c",
                    "-o", str(h), os.path.join(ROOT, "tests", "sor_schedule_host.cpp")], check=True)
    exe = tmp_path / "harness"
    subprocess.run([hipcc, "-o", str(exe), str(h)] + objs, check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    lines = out.stdout.strip().splitlines()
    assert out.returncode == 0 and lines and lines[-1] == "ALL BIT-EXACT", out.stdout[-3000:]
    assert len(lines) > 50  # every case and depth ran
