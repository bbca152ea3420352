"""The flag bits of the compact operator derived from wet depths (orctm_amd/csrc/nh_flags.h), which the
iteration kernels use in place of the flag plane when the table form is on (checked bit for bit at
create, k_zflags), compared on the CPU with the oracle's flags (orc_compact: the branches of
mo_petsc.f90:623-765 cell by cell over WETO).

* Every shared geometry: the derived plane equals the oracle's, bit for bit.
* A wet cell under a dry one (not a z-level column): the derived plane differs there, which is what
  sends such a system back to the flag plane on the device.
"""
import copy
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from cases import SPECS, build

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    so = str(tmp_path_factory.mktemp("flags") / "libflags_host.so")
    subprocess.run([gxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "flags_host.cpp")],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.derive_flag_plane.argtypes = [ctypes.POINTER(ctypes.c_double)] + [ctypes.c_int] * 5 + \
                                     [ctypes.POINTER(ctypes.c_uint16)]
    lib.derive_flag_plane.restype = None
    return lib


def derived(lib, g):
    weto = np.asfortranarray(g.weto, dtype=np.float64)
    out = np.zeros(g.cells, dtype=np.uint16)
    lib.derive_flag_plane(weto.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), g.IE, g.JE, g.KE,
                          int(g.spec.cyclic_x), int(g.spec.cyclic_y), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return out


@pytest.mark.parametrize("name", list(SPECS))
def test_derived_flags_equal_the_assembled_ones(harness, name):
    g, _ = build(name)
    _, _, fl, _ = oracle.compact(g)
    d = derived(harness, g)
    assert np.array_equal(d, fl)
    assert (fl & 1).any()
    if name == "regional":  # land: every X_C / Y_C variant and the bottom Z_C occur
        assert {1, 2, 3} <= set(np.unique((fl >> 7) & 3)) and {1, 2, 3} <= set(np.unique((fl >> 9) & 3))
        assert ((fl >> 11) & 1).any()


def test_wet_cell_under_dry_cell_is_detected(harness):
    g0, _ = build("regional")
    g = copy.copy(g0)
    weto = g0.weto.copy(order="F")
    deep = np.argwhere(g0.weto[1:-1, 1:-1, :].sum(axis=2) >= 6)[0] + 1
    i, j = int(deep[0]), int(deep[1])
    weto[i, j, 2] = 0.0  # a dry cell at level 3 above wet ones
    g.weto = weto
    g._c = None
    _, _, fl, _ = oracle.compact(g)
    d = derived(harness, g)
    bad = np.flatnonzero(d != fl)
    assert bad.size > 0
    nx, ny = g.nx, g.ny
    cols = {(int(n % nx), int((n // nx) % ny)) for n in bad}
    assert (i - 1, j - 1) in cols  # the column itself (interior index)
