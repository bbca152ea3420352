"""The serving loop's bookkeeping of one batched time step (orctm_amd/csrc/nh_serve_step.h), compiled
for the host and driven with synthetic outcomes of orctm_nh_batch_solve:

* a batch-wide failure (HIP error, out of memory, invalid batch) is no step of any member: every member
  retires, carrying that status, with its step count and iteration sums unchanged -- a request must
  never retire as OK when nothing was solved;
* otherwise each member counts the step and its iterations; it retires after its last step, or at once
  with a fatal status of its own (non-finite right-hand side), while a member that did not converge
  carries the status and goes on.
"""
import ctypes
import os
import shutil
import subprocess

import pytest

from orctm_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    so = str(tmp_path_factory.mktemp("serve") / "libserve_step.so")
    subprocess.run([gxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "serve_step_host.cpp")],
                   check=True)
    L = ctypes.CDLL(so)
    L.serve_apply_step.argtypes = [ctypes.c_int, ctypes.POINTER(abi.StatsC), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(abi.RequestStatsC),
                                   ctypes.POINTER(ctypes.c_int32)]
    L.serve_apply_step.restype = None
    return L


def step(lib, rc, stats, nsteps, rs):
    n = len(nsteps)
    st = (abi.StatsC * n)(*stats)
    ns = (ctypes.c_int32 * n)(*nsteps)
    out = (ctypes.c_int32 * n)()
    lib.serve_apply_step(rc, st, n, ns, rs, out)
    return list(out)


def stats(status=abi.OK, inner=10, outer=2, rel=1e-11):
    s = abi.StatsC()
    s.status, s.inner_iters, s.outer_iters, s.rel_res = status, inner, outer, rel
    return s


@pytest.mark.parametrize("rc", [abi.ERR_HIP, abi.ERR_NOMEM, abi.ERR_INVALID, abi.ERR_NODEVICE])
def test_batch_failure_retires_every_member_with_its_error(lib, rc):
    rs = (abi.RequestStatsC * 3)()
    rs[1].steps_done, rs[1].inner_iters = 4, 40
    # the failed solve wrote no member stats: they read as zero, i.e. status OK with rel.res 0
    out = step(lib, rc, [abi.StatsC(), abi.StatsC(), abi.StatsC()], [10, 10, 10], rs)
    assert out == [1, 1, 1]
    assert [r.status for r in rs] == [rc] * 3
    assert [r.steps_done for r in rs] == [0, 4, 0]
    assert rs[1].inner_iters == 40


def test_member_outcomes(lib):
    rs = (abi.RequestStatsC * 4)()
    rs[0].steps_done = 9  # its last step now
    sts = [stats(), stats(abi.NOT_CONVERGED, inner=100), stats(abi.ERR_NONFINITE, inner=0, outer=0), stats()]
    out = step(lib, abi.NOT_CONVERGED, sts, [10, 10, 10, 10], rs)
    assert out == [1, 0, 1, 0]
    assert [r.status for r in rs] == [abi.OK, abi.NOT_CONVERGED, abi.ERR_NONFINITE, abi.OK]
    assert [r.steps_done for r in rs] == [10, 1, 1, 1]
    assert rs[1].inner_iters == 100 and rs[3].outer_iters == 2 and rs[3].rel_res == 1e-11
    # a later good step keeps the first non-OK status
    out = step(lib, abi.OK, [stats()] * 4, [10, 10, 10, 2], rs)
    assert rs[1].status == abi.NOT_CONVERGED and rs[1].steps_done == 2
    assert out[3] == 1 and rs[3].steps_done == 2
