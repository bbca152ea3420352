"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on the same seeded inputs.

Bars: bit-exact for the assembled reference operator, the FP64 matvec, the inner matvec and the line
preconditioner (same floating-point operations in the same order); for full solves, the true relative
residual ||b - A p|| / ||b|| computed by the oracle's FP64 operator must be <= 1e-10, and the solution
must match an FP64 direct solve (SciPy SuperLU) to ||p - p_direct|| / ||p_direct|| <= 1e-7.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from cases import SPECS, build
from orctm_amd import abi, synth
from orctm_amd.nh import NHBatch, NHSolver, options

pytestmark = pytest.mark.gpu

RTOL = 1e-10
SOL_TOL = 1e-7


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    v = np.uint64 if a.dtype == np.float64 else np.uint32
    return np.array_equal(a.view(v), b.view(v))


_cache = {}


def case(name):
    if name not in _cache:
        g, divg = build(name)
        A, S = oracle.assemble(g)
        b = oracle.rhs(g, divg)
        _cache[name] = (g, divg, A, S, b)
    return _cache[name]


def direct(name):
    key = name + "/direct"
    if key not in _cache:
        g, divg, A, S, b = case(name)
        _cache[key] = spla.spsolve(oracle.to_scipy(g, A).tocsc(), b)
    return _cache[key]


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_operator_bit_exact(name, precision):
    g, divg, A, S, b = case(name)
    print(f"[operator] {name} precision={precision}", flush=True)
    with NHSolver(g, options(precision=precision)) as s:
        assert bits_equal(s.operator(), A)


@pytest.mark.parametrize("name", list(SPECS))
def test_apply64_bit_exact(name):
    g, divg, A, S, b = case(name)
    print(f"[apply64] {name}", flush=True)
    x = synth.normal(99, 1, g.cells)
    with NHSolver(g, options(precision=abi.FP64)) as s:
        assert bits_equal(s.apply(x), oracle.matvec(g, A, x))


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_inner_operator_and_line_preconditioner_bit_exact(name, precision):
    g, divg, A, S, b = case(name)
    print(f"[inner] {name} precision={precision}", flush=True)
    dt = np.float32 if precision == abi.MIXED else np.float64
    C = oracle.inner_coef(A, S, dt, grid=g)
    invD, L = oracle.line_factor(g, C)
    x = synth.normal(98, 2, g.cells).astype(dt)
    with NHSolver(g, options(precision=precision)) as s:
        assert bits_equal(s.inner_apply(x), oracle.matvec(g, C, x))
        z, rho = s.inner_precond(x)
        z_ref, rho_ref = oracle.line_apply(g, invD, L, x)
        assert bits_equal(z, z_ref)
        tol = 1e-5 if precision == abi.MIXED else 1e-12
        assert abs(rho - rho_ref) <= tol * abs(rho_ref)


@pytest.mark.parametrize("name", list(SPECS))
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_solve_matches_direct_solve(name, precision):
    g, divg, A, S, b = case(name)
    print(f"[solve] {name} precision={precision}", flush=True)
    pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
    pnh[0, :, :] = 7.0  # halo must be left untouched (mo_petsc.f90:223-227 writes the interior only)
    with NHSolver(g, options(precision=precision)) as s:
        st = s.solve(divg, pnh)
    assert st["status"] == "OK", st
    assert np.all(pnh[0, :, :] == 7.0)
    p = g.interior(pnh)
    rel = oracle.rel_res(g, A, p, b)
    assert rel <= RTOL, (rel, st)
    assert abs(rel - st["rel_res"]) <= 1e-3 * RTOL + 1e-3 * rel
    pd = direct(name)
    err = np.linalg.norm(p - pd) / np.linalg.norm(pd)
    assert err <= SOL_TOL, err


def test_mixed_and_fp64_agree_and_oracle_iteration_counts():
    g, divg, A, S, b = case("regional")
    out = {}
    for prec in (abi.FP64, abi.MIXED):
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=prec)) as s:
            st = s.solve(divg, pnh)
        out[prec] = (g.interior(pnh), st)
        p_cpu, st_cpu = oracle.solve(g, A, S, b, precision=prec)
        # same algorithm: iteration counts agree up to reduction-order effects
        assert abs(st["inner_iters"] - st_cpu.inner_iters) <= max(3, st_cpu.inner_iters // 10)
        assert st["outer_iters"] == st_cpu.outer_iters or abs(st["outer_iters"] - st_cpu.outer_iters) <= 1
    p64, pmx = out[abi.FP64][0], out[abi.MIXED][0]
    assert np.linalg.norm(pmx - p64) / np.linalg.norm(p64) <= SOL_TOL


def test_warm_start_resolve_is_idempotent():
    g, divg, A, S, b = case("regional")
    pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
    with NHSolver(g, options(precision=abi.MIXED, warm_start=True)) as s:
        st1 = s.solve(divg, pnh)
        p1 = g.interior(pnh).copy()
        st2 = s.solve(divg, pnh)  # warm start from the converged solution: nothing left to do
    assert st1["status"] == "OK" and st2["status"] == "OK"
    assert st2["inner_iters"] == 0 and st2["outer_iters"] == 0
    assert bits_equal(g.interior(pnh), p1)


def test_warm_start_after_any_number_of_refinement_steps():
    """For every case, both precisions and inner tolerances that take 1 to 5 refinement steps (both
    parities of the step count occur): get_solution returns the solved field (oracle rel.res, direct
    solve), a warm re-solve finds nothing left to do and returns the same bits, and a perturbed
    warm-started step equals the same step on a fresh handle given that solution by set_guess, bit for
    bit and with the same iteration counts."""
    parities = set()
    # looser inner tolerances take more refinement steps
    configs = [(abi.MIXED, None), (abi.MIXED, 1e-2), (abi.FP64, None), (abi.FP64, 1e-3)]
    for name in ("regional", "closed", "cyclic_x", "transect"):
        g, divg, A, S, b = case(name)
        for prec, rin in configs:
            print(f"[warm steps] {name} precision={prec} rtol_inner={rin}", flush=True)
            zero = lambda: np.zeros((g.IE, g.JE, g.KE), order="F")
            with NHSolver(g, options(precision=prec, rtol_inner=rin)) as s:
                s.set_rhs(divg)
                st1 = s.solve_resident()
                assert st1["status"] == "OK", st1
                parities.add(st1["outer_iters"] % 2)
                p1 = s.get_solution(zero())
                assert oracle.rel_res(g, A, g.interior(p1), b) <= RTOL
                pd = direct(name)
                assert np.linalg.norm(g.interior(p1) - pd) / np.linalg.norm(pd) <= SOL_TOL
                st2 = s.solve_resident()
                assert st2["inner_iters"] == 0 and bits_equal(s.get_solution(zero()), p1), st2
                s.perturb_rhs(0.01, 99)
                st3 = s.solve_resident()
                assert st3["status"] == "OK" and st3["inner_iters"] > 0, st3
                p3 = s.get_solution(zero())
            with NHSolver(g, options(precision=prec, rtol_inner=rin)) as f:
                f.set_rhs(divg)
                f.perturb_rhs(0.01, 99)
                f.set_guess(p1)
                st4 = f.solve_resident()
                p4 = f.get_solution(zero())
            assert st4["inner_iters"] == st3["inner_iters"] and st4["outer_iters"] == st3["outer_iters"]
            assert bits_equal(p3, p4), (name, prec, rin)
    assert parities == {0, 1}, parities


def test_set_guess_and_cold_start_options():
    g, divg, A, S, b = case("cyclic_x")
    pd = direct("cyclic_x")
    guess = g.to_fortran(pd)
    with NHSolver(g, options(precision=abi.FP64, warm_start=False)) as s:
        s.set_rhs(divg)
        s.set_guess(guess)
        st = s.solve_resident()
        assert st["status"] == "OK" and st["inner_iters"] == 0
        st = s.solve_resident()  # warm_start off: starts from zero again
        assert st["status"] == "OK" and st["inner_iters"] > 0


def test_zero_rhs_gives_zero_solution():
    g, divg, A, S, b = case("closed")
    pnh = np.full((g.IE, g.JE, g.KE), 3.0, order="F")
    with NHSolver(g, options(precision=abi.MIXED)) as s:
        st = s.solve(np.zeros_like(divg), pnh)
    assert st["status"] == "OK" and st["rel_res"] == 0.0
    assert np.all(g.interior(pnh) == 0.0)


def test_all_land_system():
    spec = synth.GridSpec(16, 8, 5, bathy="flat", depth=0.0, seed=3)
    g, divg = synth.build(spec)
    assert g.weto.sum() == 0
    pnh = np.ones((g.IE, g.JE, g.KE), order="F")
    with NHSolver(g, options(precision=abi.MIXED)) as s:
        st = s.solve(divg, pnh)
    assert st["status"] == "OK"
    assert np.all(g.interior(pnh) == 0.0)


def test_nonfinite_rhs_is_reported_and_pnh_untouched():
    g, divg, A, S, b = case("closed")
    bad = divg.copy(order="F")
    wet = np.argwhere(g.weto[1:-1, 1:-1, :] == 1.0)[0] + np.array([1, 1, 0])
    bad[tuple(wet)] = np.nan
    pnh = np.full((g.IE, g.JE, g.KE), 5.0, order="F")
    with NHSolver(g, options(precision=abi.MIXED)) as s:
        st = s.solve(bad, pnh, check=False)
    assert st["rc"] == abi.ERR_NONFINITE
    assert np.all(pnh == 5.0)


def test_maxit_exhaustion_returns_status():
    g, divg, A, S, b = case("transect")
    pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
    with NHSolver(g, options(precision=abi.FP64, maxit=5)) as s:
        st = s.solve(divg, pnh, check=False)
    assert st["rc"] == abi.NOT_CONVERGED
    assert st["inner_iters"] <= 5 + 16
    assert np.all(np.isfinite(pnh))


def test_invalid_grid_rejected():
    g, divg = synth.build(synth.GridSpec(12, 6, 4, seed=4))
    g.dlxp[3, 3] = 0.0
    with pytest.raises(Exception) as ei:
        NHSolver(g, options(precision=abi.MIXED))
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_batch_results_identical_alone_and_in_the_mix(precision):
    names = ["regional", "transect", "cyclic_x", "closed", "stretched_dx"]
    alone = {}
    for n in names:
        g, divg, A, S, b = case(n)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision)) as s:
            st = s.solve(divg, pnh)
        assert st["status"] == "OK"
        alone[n] = (g.interior(pnh), st)
    solvers = []
    for n in names:
        g, divg, A, S, b = case(n)
        s = NHSolver(g, options(precision=precision))
        s.set_rhs(divg)
        solvers.append(s)
    batch = NHBatch(solvers)
    rc, stats = batch.solve()
    assert rc == abi.OK, stats
    for n, s, st in zip(names, solvers, stats):
        g = s.grid
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        assert bits_equal(g.interior(pnh), alone[n][0]), n
        assert st["inner_iters"] == alone[n][1]["inner_iters"]
        assert st["outer_iters"] == alone[n][1]["outer_iters"]
    batch.close()
    for s in solvers:
        s.close()


def test_clone_shares_operator_and_solves_identically():
    g, divg, A, S, b = case("regional")
    s = NHSolver(g, options(precision=abi.MIXED))
    s.set_rhs(divg)
    c = s.clone()  # copies b, shares the operator
    st1 = s.solve_resident()
    p1 = s.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))
    s.close()  # the clone keeps the shared operator alive
    st2 = c.solve_resident()
    p2 = c.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))
    c.close()
    assert st1["rc"] == abi.OK and st2["rc"] == abi.OK
    assert bits_equal(p1, p2)


def test_perturbed_rhs_is_deterministic_and_keeps_land_dry():
    g, divg, A, S, b = case("closed")
    base = NHSolver(g, options(precision=abi.FP64))
    base.set_rhs(divg)
    sols = []
    for seed in (5, 5, 6):
        c = base.clone()
        c.perturb_rhs(0.01, seed)
        st = c.solve_resident()
        assert st["rc"] == abi.OK
        sols.append(g.interior(c.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))))
        c.close()
    base.close()
    assert bits_equal(sols[0], sols[1])
    assert not np.array_equal(sols[0], sols[2])
    # the perturbed right-hand side is b*(1+0.01*xi): zero on land, close to b elsewhere
    bp = oracle.matvec(g, A, sols[0])
    wet = g.interior(g.weto) == 1.0
    assert np.all(np.abs(bp[~wet]) <= 1e-12 * np.abs(b).max())
    rel = np.linalg.norm(bp - b) / np.linalg.norm(b)
    assert 0.002 < rel < 0.05


@pytest.mark.parametrize("name", ["regional", "stretched_dx"])
@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_tile_height_does_not_change_results(name, precision):
    """Block sums are formed over fixed 4-row groups, so every tile height gives the same bits."""
    print(f"[tile] {name} precision={precision}", flush=True)
    g, divg, A, S, b = case(name)
    out = []
    for rows in (4, 8, 12, 0):
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, tile_rows=rows)) as s:
            st = s.solve(divg, pnh)
        assert st["status"] == "OK", st
        assert st["tile_rows"] == rows or (rows == 0 and st["tile_rows"] in (4, 8, 12))
        out.append((g.interior(pnh), st))
    for p, st in out[1:]:
        assert bits_equal(p, out[0][0])
        assert st["inner_iters"] == out[0][1]["inner_iters"]
        assert st["outer_iters"] == out[0][1]["outer_iters"]


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_chunked_and_streaming_inner_kernels_agree(precision):
    """A launch of fewer tiles than CUs (256) runs the inner kernels in their chunked, latency-bound
    forms (levels in chunks, one barrier per chunk); a larger one streams a level at a time. On
    512x132x6, 4-row tiles give 8 x 33 = 264 tiles and 12-row tiles 8 x 11 = 88, so the two forms
    meet on one system: the same bits for the inner operator, the line solve and the whole solve."""
    print(f"[chunked vs streaming] precision={precision}", flush=True)
    g, divg = synth.build(synth.GridSpec(512, 132, 6, smooth=3.0, seed=synth.BASE_SEED + 77))
    x = synth.normal(5, 3, g.cells)
    res = {}
    for rows in (4, 12):
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, tile_rows=rows, warm_start=False)) as s:
            dt = np.float32 if precision == abi.MIXED else np.float64
            y = s.inner_apply(x.astype(dt))
            z, _ = s.inner_precond(x.astype(dt))
            st = s.solve(divg, pnh)
        assert st["status"] == "OK" and st["tile_rows"] == rows, st
        res[rows] = (y, z, g.interior(pnh), st)
    (y4, z4, p4, s4), (y12, z12, p12, s12) = res[4], res[12]
    A, S = oracle.assemble(g)
    dt = np.float32 if precision == abi.MIXED else np.float64
    assert bits_equal(y4, oracle.matvec(g, oracle.inner_coef(A, S, dt, grid=g), x.astype(dt)))
    assert bits_equal(y4, y12) and bits_equal(z4, z12)
    assert bits_equal(p4, p12)
    assert s4["inner_iters"] == s12["inner_iters"] and s4["outer_iters"] == s12["outer_iters"]


# Grids for the two addressing forms of the matvec kernel: small launches (deep-prefetch form), a
# periodic grid whose last tile column wraps (lanes past nx hold wrapped columns but must not store), the
# transect, and a launch of 264 tiles (streaming form at 4 rows).
ADDR_GRIDS = {
    "regional": (SPECS["regional"], 0),
    "transect": (SPECS["transect"], 0),
    "wrap": (synth.GridSpec(100, 19, 8, cyclic_x=True, cyclic_y=True, smooth=3.0, seed=synth.BASE_SEED + 79), 4),
    "streaming": (synth.GridSpec(512, 132, 6, smooth=3.0, seed=synth.BASE_SEED + 77), 4),
}


@pytest.mark.parametrize("name", list(ADDR_GRIDS))
@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
@pytest.mark.parametrize("zw_table", [-1, 1])
def test_buffer_and_pointer_addressing_agree(name, precision, zw_table):
    """The matvec kernel addresses memory through buffer descriptors (masked lanes turned away by the
    range check) or through 64-bit pointers (options.buffer_addressing = -1): the same bits for the
    inner operator (against the oracle) and the whole solve, with the ZW/ZP table and without."""
    print(f"[addressing] {name} precision={precision} zw_table={zw_table}", flush=True)
    spec, rows = ADDR_GRIDS[name]
    g, divg = synth.build(spec)
    dt = np.float32 if precision == abi.MIXED else np.float64
    x = synth.normal(9, 3, g.cells).astype(dt)
    res = {}
    for addr in (0, -1):
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, tile_rows=rows, zw_table=zw_table, buffer_addressing=addr,
                                 warm_start=False)) as s:
            y = s.inner_apply(x)
            st = s.solve(divg, pnh)
        assert st["status"] == "OK", st
        assert st["buffer_addressing"] == (1 if addr == 0 else 0), st
        res[addr] = (y, g.interior(pnh), st)
    (y0, p0, s0), (y1, p1, s1) = res[0], res[-1]
    A, S = oracle.assemble(g)
    assert bits_equal(y0, oracle.matvec(g, oracle.inner_coef(A, S, dt, grid=g), x))
    assert bits_equal(y0, y1)
    assert bits_equal(p0, p1)
    assert s0["inner_iters"] == s1["inner_iters"] and s0["outer_iters"] == s1["outer_iters"]
    assert s0["zw_table"] == s1["zw_table"]


def test_invalid_buffer_addressing_rejected():
    g, divg, A, S, b = case("closed")
    with pytest.raises(Exception) as ei:
        NHSolver(g, options(precision=abi.MIXED, buffer_addressing=1))
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID


def test_batch_with_mixed_tile_heights_matches_alone():
    names = ["regional", "cyclic_x", "one_layer"]
    alone, solvers = {}, []
    for n, rows in zip(names, (4, 8, 12)):
        g, divg, A, S, b = case(n)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=abi.MIXED, tile_rows=12)) as s:
            st = s.solve(divg, pnh)
        alone[n] = (g.interior(pnh), st)
        s = NHSolver(g, options(precision=abi.MIXED, tile_rows=rows))
        s.set_rhs(divg)
        solvers.append(s)
    batch = NHBatch(solvers)
    rc, stats = batch.solve()
    assert rc == abi.OK, stats
    assert [st["tile_rows"] for st in stats] == [4, 8, 12]
    for n, s, st in zip(names, solvers, stats):
        g = s.grid
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        assert bits_equal(g.interior(pnh), alone[n][0]), n
        assert st["inner_iters"] == alone[n][1]["inner_iters"]
    batch.close()
    for s in solvers:
        s.close()


def test_invalid_tile_rows_rejected():
    g, divg, A, S, b = case("closed")
    with pytest.raises(Exception) as ei:
        NHSolver(g, options(precision=abi.MIXED, tile_rows=6))
    assert getattr(ei.value, "code", None) == abi.ERR_INVALID


@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_batch_vertical_pairs_match_alone(precision):
    """Systems whose rows hold an odd number of tiles (300 and 140 columns: 5 and 3 tiles of 64), one of
    them with an odd number of row-blocks, batched into one launch of 391 4-row tiles: the column update
    pairs the last tiles of two row-blocks in one block (and runs the last one alone), where each system
    solved on its own takes the one-tile-per-block kernel. Same bits either way."""
    specs = [synth.GridSpec(300, 146, 12, smooth=3.0, seed=71), synth.GridSpec(140, 150, 12, smooth=3.0, seed=72),
             synth.GridSpec(200, 90, 12, smooth=3.0, seed=73)]
    grids = [synth.build(sp) for sp in specs]
    alone, solvers = [], []
    for g, divg in grids:
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, tile_rows=4)) as s:
            st = s.solve(divg, pnh)
        assert st["status"] == "OK", st
        alone.append((g.interior(pnh), st))
        s = NHSolver(g, options(precision=precision, tile_rows=4))
        s.set_rhs(divg)
        solvers.append(s)
    batch = NHBatch(solvers)
    rc, stats = batch.solve()
    assert rc == abi.OK, stats
    for k, (s, st) in enumerate(zip(solvers, stats)):
        g = s.grid
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        s.get_solution(pnh)
        assert bits_equal(g.interior(pnh), alone[k][0]), k
        assert st["inner_iters"] == alone[k][1]["inner_iters"]
        assert st["rel_res"] <= RTOL
    batch.close()
    for s in solvers:
        s.close()
