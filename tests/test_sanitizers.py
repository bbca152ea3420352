"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code the tests rely on, run on the host:

* the oracle (oracle/*.c: assembly, compact form, right-hand side, both solvers, OCVTOT_NH, the bottom
  solve, ITPREP/TRONEU), driven by tests/sanitize/oracle_driver.c on small grids with land, partial
  cells and every side type, leak detection on;
* the host-compiled library headers the tests load as harnesses (the flag derivation nh_flags.h and the
  serving loop's step bookkeeping nh_serve_step.h), driven by tests/sanitize/harness_driver.cpp;
* the thread-by-thread host run of the device's blocked SOR schedule (tests/sor_schedule_host.cpp, built
  for the host only), which also checks its results bit for bit against the oracle.
Any out-of-bounds access, use after free, leak (oracle) or undefined arithmetic aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="2")


def _run(exe, expect, timeout=600):
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0 and r.stdout.strip().splitlines()[-1] == expect, (r.stdout[-2000:] + r.stderr[-4000:])


def test_oracle_under_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    exe = str(tmp_path / "oracle_san")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c")] + \
           [os.path.join(ROOT, "oracle", f) for f in ("nh_oracle.c", "nh_oracle_ocv.c", "nh_oracle_sor.c",
                                                     "nh_oracle_bot.c")]
    subprocess.run([gcc, "-std=c11", "-fopenmp", "-ffp-contract=off", *SAN, "-o", exe, *srcs, "-lm"], check=True)
    _run(exe, "SANITIZED OK")


def test_host_harnesses_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "harness_san")
    subprocess.run([gxx, "-std=c++17", *SAN, "-o", exe, os.path.join(ROOT, "tests", "sanitize", "harness_driver.cpp")],
                   check=True)
    _run(exe, "SANITIZED OK")


def test_sor_schedule_under_sanitizers(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    gcc = shutil.which("gcc")
    if hipcc is None or gcc is None:
        pytest.skip("needs hipcc and gcc")
    objs = []
    for src in ("nh_oracle_ocv.c", "nh_oracle_sor.c"):
        o = str(tmp_path / (src + ".o"))
        subprocess.run([gcc, "-std=c11", "-ffp-contract=off", *SAN, "-c", "-o", o, os.path.join(ROOT, "oracle", src)],
                       check=True)
        objs.append(o)
    # host-only build: every sanitizer flag applies to the host pass
    host_san = []
    for f in SAN:
        host_san += ["-Xarch_host", f] if f.startswith("-fsanitize") or f.startswith("-fno-sanitize") else [f]
    h = str(tmp_path / "harness.o")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-ffp-contract=off", "-w", *host_san,
                    "-c", "-o", h, os.path.join(ROOT, "tests", "sor_schedule_host.cpp")], check=True)
    exe = str(tmp_path / "sor_san")
    link_san = [x for pair in (("-Xarch_host", "-fsanitize=address,undefined"),) for x in pair]
    subprocess.run([hipcc, *link_san, "-o", exe, h, *objs], check=True)
    _run(exe, "ALL BIT-EXACT")
