"""The Fortran binding (orctm_amd/fortran/orctm_nh_c.f90) run from a Fortran host on the GPU.

tests/fortran/driver.F90 makes ORCTM's calls -- petsc_start, petsc_solve_3D twice (the second
warm-started), ocvtot_nh_device, troneu_device, petsc_end -- with the model's arrays held in module
globals shaped as mo_commo1.f90:105-154 allocates them. The same calls are made here through the Python
mirror of the C ABI with the same inputs; every array the Fortran side produces must carry the same bits.
That checks what the binding adds: the struct layouts, the argument order, the halo'd bounds
(I_start:IE, J_start:JE, KEP), the flags from MO_OBCS and the copy of DZW. Built by
tests/fortran/build.py (also run by __graft_entry__.build()) for the "regional" geometry (KE = 13).
"""
import os
import subprocess

import numpy as np
import pytest

from fortran_host import FBUILD, bits_equal, inputs, read_output, write_input
from orctm_amd import abi
from orctm_amd.nh import NHBarotropic, NHSolver, options

pytestmark = pytest.mark.gpu


def python_path(g, divg, v, b, b1o, cg2d):
    """The Fortran host's calls through the Python mirror: petsc_start/solve_3D/ocvtot/troneu/petsc_end."""
    out = []
    with NHSolver(g, options(bottom_solve=cg2d)) as s:
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        for _ in range(2):
            st = s.solve(divg, pnh)
            assert st["rc"] == abi.OK
            out.append(pnh.copy(order="F"))
        s.attach_vgrid(v)
        d = divg.copy(order="F")
        uoo, voe, wno = (np.array(a, order="F", copy=True) for a in (v.uoo, v.voe, v.wno))
        s.divergence(v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo, d)
        bot = s.get_bottom() if cg2d else None
        assert s.solve_resident()["rc"] == abi.OK
        s.project(pnh, uoo, voe, wno)
        out += [d, pnh, uoo, voe, wno]
        if cg2d:
            out += [bot["pnh_bot"], bot["divg_bot_surf"]]
    with NHBarotropic(b) as sb:
        z1o = np.zeros((b.IE, b.JE), order="F")
        b1 = b1o.copy(order="F")
        sb.troneu(b.u1o, b.uzo, b.v1e, b.vze, b1, z1o)
        out += [z1o, b1]
    return out


@pytest.mark.parametrize("variant", ["plain", "cg2d"])
def test_fortran_host_matches_the_c_abi(variant, tmp_path):
    exe = os.path.join(FBUILD, variant, "orctm_nh_fdriver")
    if not os.path.exists(exe):
        pytest.skip("the Fortran host is not built (no amdflang when the tree was built)")
    print(f"[fortran] {variant}", flush=True)
    g, divg, v, b, b1o = inputs()
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, g, divg, v, b, b1o)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "done" in r.stdout and "status" not in r.stdout, r.stdout
    cg2d = variant == "cg2d"
    ref = python_path(g, divg, v, b, b1o, cg2d)
    got = read_output(fout, [a.shape for a in ref])
    names = ["PNH (solve 1)", "PNH (solve 2)", "DIVG", "PNH", "UOO", "VOE", "WNO"] + \
            (["PNH_bot", "DIVG_bot_surf"] if cg2d else []) + ["Z1O", "B1O"]
    for n, a, r_ in zip(names, got, ref):
        assert bits_equal(a, r_), n
    # the solves did something: the warm-started second solve agrees with the first to the tolerance
    p1, p2 = got[0], got[1]
    assert np.abs(p1).max() > 0
    assert np.linalg.norm(p2 - p1) <= 1e-8 * np.linalg.norm(p1)


def test_restart_restores_the_warm_start(tmp_path):
    """aufw_nh saves the resident pressure with the restart files and aufr_nh gives it back to a new run:
    the restarted run's next solve is bit-identical to the uninterrupted run's (which warm-starts from
    that pressure, mo_petsc.f90:232), where a restart without it starts from zero and takes a different
    path to the same tolerance. The record is a Fortran sequential header (IDATE, code 594, KE, IE*JE)
    and the (IE,JE,KE) field: the first solve's interior, zero halo."""
    exe = os.path.join(FBUILD, "restart", "orctm_nh_frestart")
    if not os.path.exists(exe):
        pytest.skip("the Fortran restart host is not built (no amdflang when the tree was built)")
    print("[fortran] restart", flush=True)
    g, divg, v, b, b1o = inputs()
    fin, frst, fout = str(tmp_path / "in.bin"), str(tmp_path / "restart.bin"), str(tmp_path / "out.bin")
    write_input(fin, g, divg, v, b, b1o)
    r = subprocess.run([exe, fin, frst, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "done" in r.stdout and "status" not in r.stdout and "aufr_nh" not in r.stdout, r.stdout
    shape = (g.IE, g.JE, g.KE)
    uninterrupted, restarted, cold = read_output(fout, [shape] * 3)
    assert bits_equal(restarted, uninterrupted)
    assert not bits_equal(cold, uninterrupted)
    assert np.linalg.norm(cold - uninterrupted) <= 1e-7 * np.linalg.norm(uninterrupted)
    raw = open(frst, "rb").read()
    n = g.IE * g.JE * g.KE
    m1 = np.frombuffer(raw[0:4], np.int32)[0]
    hdr = np.frombuffer(raw[4:20], np.int32)
    m2 = np.frombuffer(raw[20:24], np.int32)[0]
    m3 = np.frombuffer(raw[24:28], np.int32)[0]
    field = np.frombuffer(raw[28:28 + 8 * n], np.float64).reshape(shape, order="F")
    m4 = np.frombuffer(raw[28 + 8 * n:32 + 8 * n], np.int32)[0]
    assert (m1, m2, m3, m4) == (16, 16, 8 * n, 8 * n) and len(raw) == 32 + 8 * n
    assert list(hdr) == [20261203, 594, g.KE, g.IE * g.JE]
    with NHSolver(g, options()) as s:
        pnh = np.zeros(shape, order="F")
        assert s.solve(divg, pnh)["rc"] == abi.OK
    assert bits_equal(field, pnh)
