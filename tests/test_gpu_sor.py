"""GPU parity for the barotropic free-surface solve (orctm_nh_sor_create / orctm_nh_troneu: ITPREP,
itprep.f90:15-38, and TRONEU, troneu.f90:48-131) against the oracle's restatement, through the C ABI.

Bar: bit-exact Z1O and B1O after the full 250 red-black sweeps (within a colour every update reads only
the other colour, so the device's parallel order gives the sequential loop's bits).
"""
import copy

import numpy as np
import pytest

import oracle
from cases import SPECS, build
from orctm_amd import abi, synth
from orctm_amd.nh import NHBarotropic

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("name", list(SPECS))
def test_troneu_bit_exact(name):
    print(f"[sor] {name}", flush=True)
    g, divg = build(name)
    b = synth.make_barotropic(g)
    coef = oracle.itprep(b)
    b1o_in = np.asfortranarray(synth.normal(91, 1, b.IE * b.JE).reshape(b.JE, b.IE).T)
    z_ref, b1_ref = oracle.troneu(b, coef, b1o=b1o_in)
    b1o = b1o_in.copy(order="F")
    z1o = np.full((b.IE, b.JE), 3.0, order="F")
    with NHBarotropic(b) as s:
        ms = s.troneu(b.u1o, b.uzo, b.v1e, b.vze, b1o, z1o)
        assert ms > 0.0
        assert bits_equal(z1o, z_ref)
        assert bits_equal(b1o, b1_ref)
        # a second call on the same handle (the captured sweep graph is replayed)
        z2 = np.zeros_like(z1o)
        b2 = b1o_in.copy(order="F")
        s.troneu(b.u1o, b.uzo, b.v1e, b.vze, b2, z2)
        assert bits_equal(z2, z_ref)


def test_sor_rejects_bad_relaxation():
    g, divg = build("closed")
    b = copy.copy(synth.make_barotropic(g))
    b.sorpar = 2.0
    with pytest.raises(Exception) as ei:
        NHBarotropic(b)
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID
