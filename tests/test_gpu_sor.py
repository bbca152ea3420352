"""GPU parity for the barotropic free-surface solve (orctm_nh_sor_create / orctm_nh_troneu: ITPREP,
itprep.f90:15-38, and TRONEU, troneu.f90:48-131) against the oracle's restatement, through the C ABI.

Bar: bit-exact Z1O and B1O after the full 250 red-black sweeps (within a colour every update reads only
values from before its half-sweep, so the device's parallel and temporally blocked order gives the
sequential loop's bits), for every compiled number of half-sweeps per launch, on the shared geometries
and on cyclic domains with odd periods (where a cell's neighbour across the seam has its own colour).
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


ODD = {
    # doubly periodic with odd periods in both directions (IE-2 = 31, JE-2 = 17)
    "odd_cyclic_xy": synth.GridSpec(31, 17, 3, bathy="flat", depth=900.0, cyclic_x=True, cyclic_y=True, seed=21),
    # periodic x with an odd period, closed north and south
    "odd_cyclic_x": synth.GridSpec(45, 26, 4, cyclic_x=True, open_n=False, open_s=False, smooth=2.0, seed=22),
    # a domain smaller than one launch's halo
    "tiny_cyclic_xy": synth.GridSpec(5, 3, 2, bathy="flat", depth=500.0, cyclic_x=True, cyclic_y=True, seed=23),
}


def _troneu_case(g, iterations=None, half_sweeps=0):
    b = copy.copy(synth.make_barotropic(g))
    if iterations is not None:
        b.iterations = iterations
    coef = oracle.itprep(b)
    b1o_in = np.asfortranarray(synth.normal(93, 1, b.IE * b.JE).reshape(b.JE, b.IE).T)
    z_ref, b1_ref = oracle.troneu(b, coef, b1o=b1o_in)
    b1o = b1o_in.copy(order="F")
    z1o = np.full((b.IE, b.JE), -2.0, order="F")
    with NHBarotropic(b, half_sweeps=half_sweeps) as s:
        s.troneu(b.u1o, b.uzo, b.v1e, b.vze, b1o, z1o)
        info = s.info()
    return b, info, bits_equal(z1o, z_ref), bits_equal(b1o, b1_ref)


@pytest.mark.parametrize("half_sweeps", [4, 6, 7, 8, 12])
@pytest.mark.parametrize("name", ["regional", "cyclic_xy", "transect"])
def test_troneu_every_launch_depth(name, half_sweeps):
    print(f"[sor depth] {name} {half_sweeps}", flush=True)
    g, divg = build(name)
    b, info, z_ok, b_ok = _troneu_case(g, half_sweeps=half_sweeps)
    assert info["half_sweeps"] == half_sweeps
    assert info["launches"] == 1 + -(-(2 * b.iterations - 1) // half_sweeps)
    assert z_ok and b_ok


@pytest.mark.parametrize("half_sweeps", [0, 4, 6, 7, 8, 12])
@pytest.mark.parametrize("name", list(ODD))
def test_troneu_odd_cyclic_periods(name, half_sweeps):
    print(f"[sor odd] {name} {half_sweeps}", flush=True)
    g, _ = synth.build(ODD[name])
    b, info, z_ok, b_ok = _troneu_case(g, half_sweeps=half_sweeps)
    assert info["half_sweeps"] in (4, 6, 7, 8, 12)
    assert z_ok and b_ok


@pytest.mark.parametrize("iterations", [0, 1, 2, 7])
def test_troneu_few_iterations(iterations):
    print(f"[sor iterations] {iterations}", flush=True)
    g, divg = build("cyclic_x")
    b, info, z_ok, b_ok = _troneu_case(g, iterations=iterations)
    assert z_ok and b_ok
    if iterations == 0:
        assert info["launches"] == 0


def test_sor_rejects_uncompiled_depth():
    g, divg = build("closed")
    b = synth.make_barotropic(g)
    with pytest.raises(Exception) as ei:
        NHBarotropic(b, half_sweeps=5)
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID


def test_sor_rejects_bad_relaxation():
    g, divg = build("closed")
    b = copy.copy(synth.make_barotropic(g))
    b.sorpar = 2.0
    with pytest.raises(Exception) as ei:
        NHBarotropic(b)
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID
