"""The per-level ZW/ZP table with a 3-level window per column (the compact operator's vertical factors,
mo_petsc.f90:557-590): on the shared geometries, whose partial cells sit at the bottom of each column as
boden.f90:182-213 builds them, every system takes the table form (stats zw_table = 1) and stays
bit-exact (inner operator) and within the solve tolerances; a grid with a partial cell in the middle of
a column does not fit the window, falls back to the per-cell planes (zw_table = 0), and is just as exact.
"""
import copy

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from cases import SPECS, build
from orctm_amd import abi
from orctm_amd.nh import NHSolver, options

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    v = np.uint64 if a.dtype == np.float64 else np.uint32
    return np.array_equal(a.view(v), b.view(v))


def check(g, divg, precision):
    A, S = oracle.assemble(g)
    b = oracle.rhs(g, divg)
    with NHSolver(g, options(precision=precision, warm_start=False)) as s:
        dt = np.float32 if precision == abi.MIXED else np.float64
        C = oracle.inner_coef(A, S, dt, grid=g)
        x = np.random.default_rng(5).standard_normal(g.cells).astype(dt)
        assert bits_equal(s.inner_apply(x), oracle.matvec(g, C, x))
        s.set_rhs(divg)
        st = s.solve_resident()
        p = g.interior(s.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F")))
    assert st["rc"] == abi.OK
    assert oracle.rel_res(g, A, p, b) <= 1e-10
    pd = spla.spsolve(oracle.to_scipy(g, A).tocsc(), b)
    assert np.linalg.norm(p - pd) <= 1e-7 * np.linalg.norm(pd)
    return st


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
@pytest.mark.parametrize("name", list(SPECS))
def test_shared_geometries_use_the_table(name, precision):
    print(f"[ztable] {name} {precision}", flush=True)
    g, divg = build(name)
    st = check(g, divg, precision)
    assert st["zw_table"] == 1


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_mid_column_partial_cell_falls_back(precision):
    g0, divg = build("regional")
    g = copy.copy(g0)
    ddpo = g0.ddpo.copy(order="F")
    wet = np.argwhere(g0.weto[1:-1, 1:-1, :].sum(axis=2) >= 8)[0] + 1  # a column wet at least 8 levels deep
    i, j = int(wet[0]), int(wet[1])
    ddpo[i, j, 3] = 0.5 * ddpo[i, j, 3]  # a partial cell four levels above the bottom region
    g.ddpo = ddpo
    g._c = None  # the C view of the grid is built from the arrays on first use
    st = check(g, divg, precision)
    assert st["zw_table"] == 0
