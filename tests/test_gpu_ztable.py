"""The per-level ZW/ZP table with a 3-level window per column (the compact operator's vertical factors,
mo_petsc.f90:557-590): on the shared geometries, whose partial cells sit at the bottom of each column as
boden.f90:182-213 builds them, every system takes the table form when asked (options zw_table = 1, stats
zw_table = 1) and stays bit-exact (inner operator) and within the solve tolerances; a grid with a partial
cell in the middle of a column does not fit the window, falls back to the per-cell planes (zw_table = 0),
and is just as exact. The table form also derives each cell's flag bits from the wet depths of its
column and of the 4 neighbours (checked against the assembled flags at create): a wet cell under a dry
one falls back the same way. By default (zw_table = 0) the table is taken only by launches of at least
2^20 cells, and a solve gives the same bits whichever form it reads.
"""
import copy

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from cases import SPECS, build
from orctm_amd import abi, synth
from orctm_amd.nh import NHSolver, options

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    v = np.uint64 if a.dtype == np.float64 else np.uint32
    return np.array_equal(a.view(v), b.view(v))


def check(g, divg, precision, zw_table=1):
    A, S = oracle.assemble(g)
    b = oracle.rhs(g, divg)
    with NHSolver(g, options(precision=precision, warm_start=False, zw_table=zw_table)) as s:
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


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_wet_cell_under_dry_cell_falls_back(precision):
    g0, divg = build("regional")
    g = copy.copy(g0)
    weto = g0.weto.copy(order="F")
    deep = np.argwhere(g0.weto[1:-1, 1:-1, :].sum(axis=2) >= 6)[0] + 1
    i, j = int(deep[0]), int(deep[1])
    weto[i, j, 2] = 0.0  # a dry cell at level 3 above wet ones: not a z-level column
    g.weto = weto
    g._c = None
    st = check(g, divg, precision)
    assert st["zw_table"] == 0


def solve_bits(g, divg, precision, zw_table):
    with NHSolver(g, options(precision=precision, warm_start=False, zw_table=zw_table)) as s:
        s.set_rhs(divg)
        st = s.solve_resident()
        p = s.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))
    assert st["rc"] == abi.OK
    return st, p


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_choice_of_form_gives_the_same_bits(precision):
    g, divg = build("regional")
    print(f"[ztable] choice {precision} cells={g.cells}", flush=True)
    runs = {m: solve_bits(g, divg, precision, m) for m in (1, 0, -1)}
    assert runs[1][0]["zw_table"] == 1
    assert runs[0][0]["zw_table"] == 0 and runs[-1][0]["zw_table"] == 0  # a small system: planes by default
    assert bits_equal(runs[1][1], runs[-1][1]) and bits_equal(runs[0][1], runs[-1][1])
    assert runs[1][0]["inner_iters"] == runs[-1][0]["inner_iters"]


def test_default_takes_the_table_at_2_20_cells():
    g, divg = synth.build(synth.GridSpec(nx=160, ny=128, nz=52, seed=synth.BASE_SEED + 71))
    assert g.nx * g.ny * g.nz >= 1 << 20
    print("[ztable] default nx=160 ny=128 nz=52", flush=True)
    st0, p0 = solve_bits(g, divg, abi.MIXED, 0)
    st1, p1 = solve_bits(g, divg, abi.MIXED, -1)
    assert st0["zw_table"] == 1 and st1["zw_table"] == 0
    assert bits_equal(p0, p1)


def test_zw_table_option_range():
    g, divg = build("closed")
    for bad in (2, -2):
        with pytest.raises(Exception) as ei:
            NHSolver(g, options(zw_table=bad))
        assert getattr(ei.value, "code", None) == abi.ERR_INVALID
