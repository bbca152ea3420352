"""GPU parity at the benchmarked workloads: the full-size systems of the five workload recipes, through
the C ABI, against the CPU oracle on the same seeded inputs.

The oracle's reference solution is its FP64 PCG solve (oracle/nh_oracle.c, the line-preconditioned CG
that stands in for the reference's KSPSolve of mo_petsc.f90:211-233) taken to rel.res 1e-12, tighter than
the device's 1e-10, so that its own error is small beside the bound checked. Bars:
  * rel.res ||b - A p|| / ||b|| of the device solution, computed by the oracle's FP64 operator: <= 1e-10;
  * ||p - p_oracle|| / ||p_oracle|| <= 1e-7 (the CPU oracle shows 2e-9 for FP64 PCG at 1e-10 and 5e-9 for
    the mixed restatement on the 1024x768x60 tile, and 1.6e-9 / 2.4e-9 on the 512x4x100 transect);
  * the forms of the inner kernels a launch can take (tile heights 4, 8, 12; streaming and chunked; the
    ZW/ZP table and the per-cell planes) give the same bits.
The systems are large (up to 252 M cells); each test states the seconds it may take.
"""
import numpy as np
import pytest

import oracle
from orctm_amd import abi, synth
from orctm_amd.nh import NHBatch, NHServer, NHSolver, options

pytestmark = pytest.mark.gpu

RTOL = 1e-10
SOL_TOL = 1e-7

_cache = {}


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    v = np.uint64 if a.dtype == np.float64 else np.uint32
    return np.array_equal(a.view(v), b.view(v))


def full(name):
    """Grid, DIVG, oracle operator, right-hand side and the oracle's FP64 solution to rel.res 1e-12."""
    if name not in _cache:
        spec = {"config1": synth.config1, "config2": synth.config2}[name]()
        g, divg = synth.build(spec)
        A, S = oracle.assemble(g)
        b = oracle.rhs(g, divg)
        p_ref, st = oracle.solve(g, A, S, b, precision=abi.FP64, rtol=1e-12)
        assert st.status == 0 and st.rel_res <= 1e-12, st
        _cache[name] = (g, divg, A, S, b, p_ref)
    return _cache[name]


def solve_on_device(g, divg, **kw):
    pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
    with NHSolver(g, options(warm_start=False, **kw)) as s:
        st = s.solve(divg, pnh)
    return g.interior(pnh), st


def check_against_oracle(tag, g, A, b, p_ref, p, st):
    rel = oracle.rel_res(g, A, p, b)
    err = float(np.linalg.norm(p - p_ref) / np.linalg.norm(p_ref))
    print(f"[{tag}] {st['inner_iters']} inner / {st['outer_iters']} outer, tile_rows {st['tile_rows']}, "
          f"zw_table {st['zw_table']}, oracle rel.res {rel:.2e}, |p - p_oracle|/|p_oracle| {err:.2e}", flush=True)
    assert st["status"] == "OK", st
    assert rel <= RTOL, (rel, st)
    assert abs(rel - st["rel_res"]) <= 1e-3 * RTOL + 1e-2 * rel  # the device's own FP64 residual agrees
    assert err <= SOL_TOL, err
    return rel, err


@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_config1_transect_512x4x100(precision):
    """The ISW transect at its full size (periodic y, closed x ends, Gaussian ridge): 2048 columns make
    launches of fewer tiles than CUs, so the inner kernels take their chunked / deep-prefetch forms."""
    g, divg, A, S, b, p_ref = full("config1")
    assert (g.nx, g.ny, g.nz) == (512, 4, 100)
    p, st = solve_on_device(g, divg, precision=precision)
    check_against_oracle(f"config1 precision={precision}", g, A, b, p_ref, p, st)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_config2_3_regional_1024x768x60_every_tile_height(precision):
    """The regional tile at its full size in both precisions and at every tile height the tuner can pick
    (the benchmarked instantiations among them: 12-row FP32 streaming tiles, 8-row FP64 tiles with
    4-level prefetch, the paired column update), each against the oracle, and all the same bits."""
    g, divg, A, S, b, p_ref = full("config2")
    assert (g.nx, g.ny, g.nz) == (1024, 768, 60)
    res = {}
    for rows in (0, 4, 8, 12):
        p, st = solve_on_device(g, divg, precision=precision, tile_rows=rows)
        assert st["zw_table"] == 1, st  # 47 M cells: the table form, as benchmarked
        if rows:
            assert st["tile_rows"] == rows, st
        res[rows] = (p, st)
    check_against_oracle(f"config2 precision={precision} tuned", g, A, b, p_ref, *res[0])
    for rows in (4, 8, 12):
        p, st = res[rows]
        assert bits_equal(p, res[0][0]), rows
        assert st["inner_iters"] == res[0][1]["inner_iters"] and st["outer_iters"] == res[0][1]["outer_iters"]
    _cache[f"config2/p/{precision}"] = res[0][0]


@pytest.mark.timeout(300)
def test_config3_mixed_against_fp64():
    """||p_mixed - p_FP64|| / ||p_FP64|| on the regional tile, both solved on the device to 1e-10."""
    g, divg, A, S, b, p_ref = full("config2")
    p = {}
    for prec in (abi.FP64, abi.MIXED):
        key = f"config2/p/{prec}"
        if key not in _cache:
            _cache[key] = solve_on_device(g, divg, precision=prec)[0]
        p[prec] = _cache[key]
    d = float(np.linalg.norm(p[abi.MIXED] - p[abi.FP64]) / np.linalg.norm(p[abi.FP64]))
    print(f"[config3] |p_mixed - p_fp64| / |p_fp64| = {d:.2e}", flush=True)
    assert d <= SOL_TOL


def config4_systems():
    if "config4" not in _cache:
        _cache["config4"] = [synth.build(sp) for sp in synth.config4_sizes()]
    return _cache["config4"]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_config4_batch_of_64_uneven_systems(precision):
    """The 64-system batch (nx, ny ~ U{96..512}, nz ~ U{30..60}, 301 M cells) in one batched solve:
    every member's oracle rel.res <= 1e-10, and a sample of members (the smallest, the largest and two
    others) bit-identical, with the same iteration counts, to the same system solved alone."""
    systems = config4_systems()
    assert len(systems) == 64
    solvers = []
    for g, divg in systems:
        s = NHSolver(g, options(precision=precision, warm_start=False))
        s.set_rhs(divg)
        solvers.append(s)
    batch = NHBatch(solvers)
    rc, stats = batch.solve()
    batch.close()
    sols = []
    for (g, _), s in zip(systems, solvers):
        sols.append(g.interior(s.get_solution(np.zeros((g.IE, g.JE, g.KE), order="F"))))
        s.close()
    assert rc == abi.OK, [st for st in stats if st["status"] != "OK"][:3]
    worst = 0.0
    for q, ((g, divg), p, st) in enumerate(zip(systems, sols, stats)):
        A, S = oracle.assemble(g)
        rel = oracle.rel_res(g, A, p, oracle.rhs(g, divg))
        assert st["status"] == "OK" and rel <= RTOL, (q, rel, st)
        worst = max(worst, rel)
        del A, S
    cells = [g.cells for g, _ in systems]
    sample = sorted({int(np.argmin(cells)), int(np.argmax(cells)), 17, 42})
    for q in sample:
        g, divg = systems[q]
        p, st = solve_on_device(g, divg, precision=precision)
        assert bits_equal(p, sols[q]), q
        assert st["inner_iters"] == stats[q]["inner_iters"] and st["outer_iters"] == stats[q]["outer_iters"], q
    print(f"[config4 precision={precision}] 64 systems, worst oracle rel.res {worst:.2e}, "
          f"alone == in the batch for systems {sample}", flush=True)


@pytest.mark.timeout(900)
def test_config5_largest_request_in_a_mix():
    """The largest serving size, 2048x1536x80 (252 M cells, mixed), as requests of the serving loop
    next to small ones: a one-step request's solution passes the oracle's rel.res <= 1e-10 and equals
    the same solve alone bit for bit; a three-step request (right-hand side perturbed by 1 % per step,
    warm-started from the resident pressure) equals the same steps run by hand on one clone."""
    nx, ny, nz = synth.config5_size(1.0)
    assert (nx, ny, nz) == (2048, 1536, 80)
    g, divg = synth.build(synth.GridSpec(nx, ny, nz, seed=synth.BASE_SEED + 555))
    small = [synth.build(synth.GridSpec(*synth.config5_size(t), seed=synth.BASE_SEED + 556 + k))
             for k, t in enumerate((0.0, 0.3))]
    big = NHSolver(g, options(precision=abi.MIXED, warm_start=True))
    big.set_rhs(divg)
    models = []
    for gs, ds in small:
        m = NHSolver(gs, options(precision=abi.MIXED, warm_start=True))
        m.set_rhs(ds)
        models.append(m)
    out1 = np.zeros((g.IE, g.JE, g.KE), order="F")
    out3 = np.zeros((g.IE, g.JE, g.KE), order="F")
    outs = [np.zeros((gs.IE, gs.JE, gs.KE), order="F") for gs, _ in small]
    with NHServer() as srv:
        tickets = [srv.submit(models[0], 10, 0.01, 300, outs[0]),
                   srv.submit(big, 1, 0.01, 100, out1),
                   srv.submit(models[1], 10, 0.01, 400, outs[1]),
                   srv.submit(big, 3, 0.01, 200, out3)]
        stats = [srv.wait(t) for t in tickets]
    assert all(s["status"] == "OK" for s in stats), stats
    assert [s["steps_done"] for s in stats] == [10, 1, 10, 3]

    def by_hand(model, nsteps, seed):
        c = model.clone()
        for step in range(nsteps):
            st = c.solve_resident()
            assert st["rc"] == abi.OK, st
            if step + 1 < nsteps:
                c.perturb_rhs(0.01, seed + step + 1)
        p = c.get_solution(np.zeros(model.shape, order="F"))
        c.close()
        return p

    assert bits_equal(out1, by_hand(big, 1, 100))
    assert bits_equal(out3, by_hand(big, 3, 200))
    assert bits_equal(outs[0], by_hand(models[0], 10, 300))
    big.close()
    for m in models:
        m.close()
    A, S = oracle.assemble(g)
    rel = oracle.rel_res(g, A, g.interior(out1), oracle.rhs(g, divg))
    print(f"[config5] 2048x1536x80 request: device rel.res {stats[1]['rel_res']:.2e}, oracle rel.res {rel:.2e}, "
          f"{stats[1]['inner_iters']} inner iterations", flush=True)
    assert rel <= RTOL


@pytest.mark.timeout(300)
@pytest.mark.parametrize("nx", [1024, 960])
@pytest.mark.parametrize("zw_table", [1, -1])
@pytest.mark.parametrize("precision", [abi.MIXED, abi.FP64])
def test_wide_shallow_grid_runs_every_streaming_form(precision, zw_table, nx):
    """1024x220x4 gives 16 x 55 = 880 tiles at 4 rows, 16 x 28 = 448 at 8 rows and 16 x 19 = 304 at
    12 rows: every height streams (256 tiles or more) with the paired column update, as the full-size
    launches do, on a grid small enough for exact checks. 960 columns make 15 tiles across (825, 420 and
    285 tiles): the last tile of each row has no partner, and half the lanes of its paired block idle.
    The inner operator (against the oracle's inner matvec), the line solve and the whole solve are the
    same bits at every height, with the ZW/ZP table forced on (and the FP64 flag bits derived from the
    wet depths) and with the per-cell planes."""
    g, divg = synth.build(synth.GridSpec(nx, 220, 4, dz0=20.0, dz1=60.0, depth=160.0, dmin=10.0, dmax=160.0,
                                         smooth=6.0, seed=synth.BASE_SEED + 91))
    dt = np.float32 if precision == abi.MIXED else np.float64
    x = synth.normal(31, 4, g.cells).astype(dt)
    A, S = oracle.assemble(g)
    C = oracle.inner_coef(A, S, dt, grid=g)
    invD, L = oracle.line_factor(g, C)
    y_ref = oracle.matvec(g, C, x)
    z_ref, _ = oracle.line_apply(g, invD, L, x)
    res = {}
    for rows in (4, 8, 12):
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, tile_rows=rows, zw_table=zw_table, warm_start=False)) as s:
            y = s.inner_apply(x)
            z, _ = s.inner_precond(x)
            st = s.solve(divg, pnh)
        assert st["status"] == "OK" and st["tile_rows"] == rows, st
        assert st["zw_table"] == (1 if zw_table == 1 else 0), st
        assert bits_equal(y, y_ref), rows
        assert bits_equal(z, z_ref), rows
        res[rows] = (g.interior(pnh), st)
    b = oracle.rhs(g, divg)
    assert oracle.rel_res(g, A, res[4][0], b) <= RTOL
    for rows in (8, 12):
        assert bits_equal(res[rows][0], res[4][0]), rows
        assert res[rows][1]["inner_iters"] == res[4][1]["inner_iters"]


@pytest.mark.timeout(300)
def test_arrays_over_2gib_take_the_pointer_matvec():
    """A system whose FP64 inner arrays exceed 2^31 bytes (2048x2048x65, flat bottom: 272 M cells, 2.18 GB
    per array) runs the matvec kernel through 64-bit pointers (stats.buffer_addressing 0), where the
    buffer descriptors' 32-bit offsets would not reach; in mixed precision its FP32 arrays (1.09 GB) keep
    the descriptors. A few inner iterations in each precision: the residual falls and stays finite."""
    spec = synth.GridSpec(2048, 2048, 65, bathy="flat", depth=4000.0, seed=synth.BASE_SEED + 93)
    g = synth.make_grid(spec)
    divg = np.asfortranarray(g.weto * np.sin(np.arange(1, g.KE + 1, dtype=np.float64))[None, None, :])
    for precision, addr in ((abi.FP64, 0), (abi.MIXED, 1)):
        print(f"[over 2 GiB] precision={precision}", flush=True)
        pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
        with NHSolver(g, options(precision=precision, maxit=8, max_outer=1, warm_start=False)) as s:
            st = s.solve(divg, pnh, check=False)
        assert st["buffer_addressing"] == addr, st
        assert st["status"] in ("OK", "NOT_CONVERGED"), st
        assert 0.0 < st["rel_res"] < 1.0, st
        assert np.isfinite(pnh).all()
