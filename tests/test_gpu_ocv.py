"""GPU parity for the steps of OCVTOT_NH around the solve (orctm_nh_divergence / orctm_nh_project,
ocvtot_nh.f90:11-62 and 151-230) against the oracle's restatement, through the C ABI.

Bar: bit-exact (FP64, same operations in the same order) for DIVG, for the right-hand side it yields
(the solve from it is bit-identical to the solve from the oracle's DIVG), and for PNH, UOO, VOE and WNO
after the projection, halos included.
"""
import numpy as np
import pytest

import oracle
from cases import SPECS, build
from orctm_amd import abi, synth
from orctm_amd.nh import NHSolver, options

pytestmark = pytest.mark.gpu

_cache = {}


def case(name):
    if name not in _cache:
        g, divg = build(name)
        _cache[name] = (g, synth.make_velocity(g))
    return _cache[name]


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("name", list(SPECS))
def test_divergence_and_rhs_bit_exact(name):
    g, v = case(name)
    print(f"[ocv divergence] {name}", flush=True)
    ref = oracle.divergence(g, v)
    divg = np.full((g.IE, g.JE, g.KE), 7.0, order="F")
    with NHSolver(g, options(precision=abi.MIXED)) as s, NHSolver(g, options(precision=abi.MIXED)) as t:
        s.attach_vgrid(v)
        s.divergence(v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo, divg)
        assert bits_equal(divg[1:-1, 1:-1, :], ref[1:-1, 1:-1, :])
        assert np.all(divg[0, :, :] == 7.0) and np.all(divg[:, -1, :] == 7.0)  # halo untouched
        st1 = s.solve_resident()
        p1 = s.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))
        t.set_rhs(np.asfortranarray(ref))
        st2 = t.solve_resident()
        p2 = t.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))
    assert st1["rc"] == abi.OK and st2["rc"] == abi.OK
    assert st1["bnorm"] == st2["bnorm"]
    assert bits_equal(p1, p2)


@pytest.mark.parametrize("name", list(SPECS))
def test_projection_bit_exact(name):
    g, v = case(name)
    print(f"[ocv project] {name}", flush=True)
    with NHSolver(g, options(precision=abi.FP64)) as s:
        s.attach_vgrid(v)
        s.divergence(v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo)
        st = s.solve_resident()
        assert st["rc"] == abi.OK
        # the caller's PNH: halo from an earlier step (a pattern), interior to be replaced by the solution
        pnh = np.asfortranarray(synth.normal(77, 3, g.IE * g.JE * g.KE).reshape(g.KE, g.JE, g.IE).transpose(2, 1, 0))
        sol = s.get_solution(pnh.copy(order="F"))
        uoo, voe = v.uoo.copy(order="F"), v.voe.copy(order="F")
        wno = np.full((g.IE, g.JE, g.KE + 1), 5.0, order="F")
        s.project(pnh, uoo, voe, wno)
    p_ref, u_ref, v_ref, w_ref = oracle.project(g, v, sol)
    assert bits_equal(pnh, p_ref)
    assert bits_equal(uoo, u_ref)
    assert bits_equal(voe, v_ref)
    assert bits_equal(wno, w_ref)


def test_divergence_needs_attached_metrics():
    g, v = case("closed")
    with NHSolver(g, options(precision=abi.MIXED)) as s:
        with pytest.raises(Exception) as ei:
            s.divergence(v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo)
        assert getattr(ei.value, "code", None) == abi.ERR_INVALID
