"""CPU oracle for the nonhydrostatic pressure Poisson path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
package, and only to check (or time beside) the HIP library; the product path never does.
The arithmetic lives in ``nh_oracle.c`` (built to ``liboracle_nh.so`` by ``oracle/Makefile``); each
function there cites the reference file:line it restates. This module only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_nh.so")
NC = 7

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        from orctm_amd.abi import GridC, StatsC

        i, sz, vp = ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p
        L.orc_assemble.argtypes = [ctypes.POINTER(GridC), _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.POINTER(GridC), _dp, _dp]
        L.orc_scatter_interior.argtypes = [ctypes.POINTER(GridC), _dp, _dp]
        L.orc_matvec64.argtypes = [i, i, i, i, i, _dp, _dp, _dp]
        L.orc_matvec32.argtypes = [i, i, i, i, i, _fp, _fp, _fp]
        L.orc_inner_coef64.argtypes = [sz, _dp, _dp, _dp]
        L.orc_inner_coef32.argtypes = [sz, _dp, _dp, _fp]
        L.orc_line_factor64.argtypes = [i, i, i, _dp, _dp, _dp]
        L.orc_line_factor32.argtypes = [i, i, i, _fp, _fp, _fp]
        L.orc_line_apply64.argtypes = [i, i, i, _dp, _dp, _dp, _dp, _dp]
        L.orc_line_apply32.argtypes = [i, i, i, _fp, _fp, _fp, _fp, _dp]
        L.orc_pcg64.argtypes = [i, i, i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, i,
                                ctypes.POINTER(ctypes.c_int)]
        L.orc_pcg32.argtypes = [i, i, i, i, i, _fp, _fp, _fp, _fp, _fp, _fp, ctypes.c_float, i,
                                ctypes.POINTER(ctypes.c_int)]
        L.orc_solve.argtypes = [i, i, i, i, i, _dp, _dp, _dp, _dp, i, ctypes.c_double, ctypes.c_double, i, i,
                                ctypes.POINTER(StatsC), vp]
        u16p = ctypes.POINTER(ctypes.c_uint16)
        L.orc_compact.argtypes = [ctypes.POINTER(GridC), _dp, _dp, u16p, _dp]
        L.orc_rebuild64.argtypes = [i, i, i, _dp, _dp, u16p, _dp, _dp]
        L.orc_inner_coef32_compact.argtypes = [i, i, i, _dp, _dp, u16p, _dp, _fp]
        L.orc_fgmres.argtypes = [i, i, i, i, i, _dp, _dp, _dp, _dp, i, ctypes.c_double, i, ctypes.POINTER(StatsC)]
        L.orc_num_threads.restype = i
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


def _f(a):
    assert a.dtype == np.float32
    return a.ctypes.data_as(_fp)


def assemble(grid):
    """(A[7,N], S[N]) of compute_matrix (mo_petsc.f90:543-782) for a synth.Grid."""
    N = grid.cells
    A = np.zeros((NC, N))
    S = np.zeros(N)
    rc = lib().orc_assemble(ctypes.byref(grid.c_struct()), _d(A), _d(S))
    if rc != 0:
        raise ValueError(f"orc_assemble status {rc}")
    return A, S


def compact(grid):
    """(zw[(nz+1)*P], zp[nz*P], flags[nz*P] uint16, col[7*P]) -- the operator's compact form."""
    P = grid.nx * grid.ny
    zw = np.zeros((grid.nz + 1) * P)
    zp = np.zeros(grid.nz * P)
    fl = np.zeros(grid.nz * P, dtype=np.uint16)
    col = np.zeros(7 * P)
    rc = lib().orc_compact(ctypes.byref(grid.c_struct()), _d(zw), _d(zp),
                           fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _d(col))
    if rc != 0:
        raise ValueError(f"orc_compact status {rc}")
    return zw, zp, fl, col


def rebuild64(grid, zw, zp, fl, col):
    A = np.zeros((NC, grid.cells))
    lib().orc_rebuild64(grid.nx, grid.ny, grid.nz, _d(zw), _d(zp), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                        _d(col), _d(A))
    return A


def inner_coef32(grid):
    """FP32 inner operator S*(-A) evaluated in float from the compact form (the device's arithmetic)."""
    zw, zp, fl, col = compact(grid)
    C = np.zeros((NC, grid.cells), dtype=np.float32)
    lib().orc_inner_coef32_compact(grid.nx, grid.ny, grid.nz, _d(zw), _d(zp),
                                   fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _d(col), _f(C))
    return C


def rhs(grid, divg):
    """b = DIVG*WETO on the interior (mo_petsc.f90:376), natural order."""
    b = np.zeros(grid.cells)
    divg = np.asfortranarray(divg, dtype=np.float64)
    lib().orc_rhs(ctypes.byref(grid.c_struct()), _d(divg), _d(b))
    return b


def cyc(grid):
    return int(grid.spec.cyclic_x), int(grid.spec.cyclic_y)


def matvec(grid, A, x):
    A = np.ascontiguousarray(A)
    x = np.ascontiguousarray(x)
    if A.dtype == np.float32:
        y = np.zeros(grid.cells, dtype=np.float32)
        lib().orc_matvec32(grid.nx, grid.ny, grid.nz, *cyc(grid), _f(A), _f(x.astype(np.float32)), _f(y))
    else:
        y = np.zeros(grid.cells)
        lib().orc_matvec64(grid.nx, grid.ny, grid.nz, *cyc(grid), _d(A), _d(x), _d(y))
    return y


def inner_coef(A, S, dtype=np.float64, grid=None):
    """Inner operator C = S*(-A): FP64 from the assembled A and S; FP32 from the compact form of
    `grid` in float arithmetic (as the device evaluates it), or from A and S if no grid is given."""
    N = S.size
    if dtype == np.float32 and grid is not None:
        return inner_coef32(grid)
    if dtype == np.float32:
        C = np.zeros((NC, N), dtype=np.float32)
        lib().orc_inner_coef32(N, _d(np.ascontiguousarray(A)), _d(S), _f(C))
    else:
        C = np.zeros((NC, N))
        lib().orc_inner_coef64(N, _d(np.ascontiguousarray(A)), _d(S), _d(C))
    return C


def line_factor(grid, C):
    invD = np.zeros(grid.cells, dtype=C.dtype)
    L = np.zeros(grid.cells, dtype=C.dtype)
    if C.dtype == np.float32:
        lib().orc_line_factor32(grid.nx, grid.ny, grid.nz, _f(C), _f(invD), _f(L))
    else:
        lib().orc_line_factor64(grid.nx, grid.ny, grid.nz, _d(C), _d(invD), _d(L))
    return invD, L


def line_apply(grid, invD, L, r):
    z = np.zeros(grid.cells, dtype=invD.dtype)
    rho = ctypes.c_double(0.0)
    r = np.ascontiguousarray(r, dtype=invD.dtype)
    if invD.dtype == np.float32:
        lib().orc_line_apply32(grid.nx, grid.ny, grid.nz, _f(invD), _f(L), _f(r), _f(z), ctypes.byref(rho))
    else:
        lib().orc_line_apply64(grid.nx, grid.ny, grid.nz, _d(invD), _d(L), _d(r), _d(z), ctypes.byref(rho))
    return z, rho.value


def solve(grid, A, S, b, precision=0, rtol=1e-10, rtol_inner=None, maxit=10000, max_outer=40, x0=None):
    """FP64 refinement around the inner PCG (the HIP library's algorithm) on the CPU."""
    from orctm_amd.abi import StatsC

    if rtol_inner is None:
        rtol_inner = 1e-5 if precision == 1 else 0.0
    p = np.zeros(grid.cells) if x0 is None else np.array(x0, dtype=np.float64)
    st = StatsC()
    Cin = inner_coef32(grid) if precision == 1 else None
    lib().orc_solve(grid.nx, grid.ny, grid.nz, *cyc(grid), _d(np.ascontiguousarray(A)), _d(S),
                    _d(np.ascontiguousarray(b)), _d(p), precision, rtol, rtol_inner, maxit, max_outer,
                    ctypes.byref(st), Cin.ctypes.data if Cin is not None else None)
    return p, st


def fgmres(grid, A, S, b, restart=30, rtol=1e-10, maxit=20000, x0=None):
    """FGMRES(restart), the reference's KSP type, FP64, line-preconditioned."""
    from orctm_amd.abi import StatsC

    x = np.zeros(grid.cells) if x0 is None else np.array(x0, dtype=np.float64)
    st = StatsC()
    lib().orc_fgmres(grid.nx, grid.ny, grid.nz, *cyc(grid), _d(np.ascontiguousarray(A)), _d(S),
                     _d(np.ascontiguousarray(b)), _d(x), restart, rtol, maxit, ctypes.byref(st))
    return x, st


def rel_res(grid, A, p, b):
    r = b - matvec(grid, A, p)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(r) / nb) if nb > 0 else float(np.linalg.norm(r))


def to_scipy(grid, A):
    """Sparse CSR matrix of the assembled operator (for direct solves in tests)."""
    import scipy.sparse as sp

    nx, ny, nz = grid.nx, grid.ny, grid.nz
    N = grid.cells
    n = np.arange(N)
    i = n % nx
    j = (n // nx) % ny
    k = n // (nx * ny)
    cx, cy = cyc(grid)
    rows, cols, vals = [n], [n], [A[0]]
    nbrs = [
        (1, np.where(i > 0, n - 1, np.where(cx, n + nx - 1, -1))),
        (2, np.where(i < nx - 1, n + 1, np.where(cx, n - (nx - 1), -1))),
        (3, np.where(j > 0, n - nx, np.where(cy, n + nx * (ny - 1), -1))),
        (4, np.where(j < ny - 1, n + nx, np.where(cy, n - nx * (ny - 1), -1))),
        (5, np.where(k > 0, n - nx * ny, -1)),
        (6, np.where(k < nz - 1, n + nx * ny, -1)),
    ]
    for q, c in nbrs:
        m = (c >= 0) & (A[q] != 0)
        rows.append(n[m])
        cols.append(c[m])
        vals.append(A[q][m])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def num_threads() -> int:
    return lib().orc_num_threads()


# ---- OCVTOT_NH around the solve (nh_oracle_ocv.c) ------------------------------------------------


class _VGrid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("IE", "JE", "KE", "i_start", "j_start", "cyc_x", "cyc_y", "open_w",
                                            "open_e", "open_n", "open_s")] + \
               [(n, _dp) for n in ("dlxp", "dlyp", "ddpo", "weto", "amsuo", "dduo", "dtdxuo", "dlyu", "amsue",
                                   "ddue", "dpye", "dlxv", "dtdxpo", "dtdyo")] + \
               [("dt", ctypes.c_double), ("dti", ctypes.c_double)]


def _vgrid(grid, v):
    s = grid.spec
    g = _VGrid(v.IE, v.JE, v.KE, v.i_start, v.j_start, int(s.cyclic_x), int(s.cyclic_y), int(s.open_w),
               int(s.open_e), int(s.open_n), int(s.open_s))
    for n in ("dlxp", "dlyp", "ddpo", "weto", "amsuo", "dduo", "dtdxuo", "dlyu", "amsue", "ddue", "dpye", "dlxv",
              "dtdxpo", "dtdyo"):
        a = getattr(v, n)
        assert a.flags.f_contiguous and a.dtype == np.float64
        setattr(g, n, a.ctypes.data_as(_dp))
    g.dt, g.dti = v.dt, v.dti
    g._keep = v
    return g


def _ocv_lib():
    L = lib()
    if not getattr(L, "_ocv", False):
        L.orc_divergence.argtypes = [ctypes.POINTER(_VGrid)] + [_dp] * 8
        L.orc_project.argtypes = [ctypes.POINTER(_VGrid)] + [_dp] * 4
        L.orc_bounds_exch.argtypes = [_dp] + [ctypes.c_int] * 7
        L._ocv = True
    return L


def divergence(grid, v, uoo=None, voe=None, wno=None, divg=None):
    """DIVG (IE,JE,KE) of ocvtot_nh.f90:11-62 on the interior cells (halo of `divg` kept, zero if new)."""
    L = _ocv_lib()
    g = _vgrid(grid, v)
    uoo = v.uoo if uoo is None else uoo
    voe = v.voe if voe is None else voe
    wno = v.wno if wno is None else wno
    out = np.zeros((v.IE, v.JE, v.KE), order="F") if divg is None else divg
    L.orc_divergence(ctypes.byref(g), _d(uoo), _d(v.uoo_1), _d(voe), _d(v.voe_1), _d(wno), _d(v.wno_1), _d(v.zo),
                     _d(out))
    return out


def project(grid, v, pnh, uoo=None, voe=None):
    """ocvtot_nh.f90:151-236 in place on copies: returns (pnh, uoo, voe, wno) after the halo exchange and
    open-boundary copies of PNH, the U/V pressure correction and the continuity integration of WNO."""
    L = _ocv_lib()
    g = _vgrid(grid, v)
    p = np.array(pnh, order="F", dtype=np.float64, copy=True)
    u = np.array(v.uoo if uoo is None else uoo, order="F", dtype=np.float64, copy=True)
    w_ = np.array(v.voe if voe is None else voe, order="F", dtype=np.float64, copy=True)
    wn = np.zeros((v.IE, v.JE, v.KE + 1), order="F")
    L.orc_project(ctypes.byref(g), _d(p), _d(u), _d(w_), _d(wn))
    return p, u, w_, wn


# ---- barotropic SOR: ITPREP / TRONEU (nh_oracle_sor.c) -------------------------------------------


class _SorGrid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("IE", "JE", "i_start", "j_start", "cyc_x", "cyc_y", "iterations")] + \
               [(n, _dp) for n in ("dlxp", "dlyp", "dtdxpo", "dtdyo", "weto1", "dlxu", "dlyu", "deuto", "dlxv",
                                   "dlyv", "deute")] + \
               [(n, ctypes.c_double) for n in ("g", "conn", "cono", "stabn", "dt", "sorpar")]


def _sorgrid(b):
    g = _SorGrid(b.IE, b.JE, b.i_start, b.j_start, int(b.cyc_x), int(b.cyc_y), int(b.iterations))
    for n in ("dlxp", "dlyp", "dtdxpo", "dtdyo", "weto1", "dlxu", "dlyu", "deuto", "dlxv", "dlyv", "deute"):
        a = getattr(b, n)
        assert a.flags.f_contiguous and a.dtype == np.float64, n
        setattr(g, n, a.ctypes.data_as(_dp))
    for n in ("g", "conn", "cono", "stabn", "dt", "sorpar"):
        setattr(g, n, float(getattr(b, n)))
    g._keep = b
    return g


def _sor_lib():
    L = lib()
    if not getattr(L, "_sor", False):
        L.orc_itprep.argtypes = [ctypes.POINTER(_SorGrid)] + [_dp] * 4
        L.orc_troneu.argtypes = [ctypes.POINTER(_SorGrid)] + [_dp] * 10
        L._sor = True
    return L


def itprep(b):
    """XX (IE,JE), UF (I_start:IE,JE), VF (IE,J_start:JE), FF (IE,JE) of itprep.f90:15-38."""
    L = _sor_lib()
    g = _sorgrid(b)
    IE, JE = b.IE, b.JE
    xx = np.zeros((IE, JE), order="F")
    uf = np.zeros((IE - b.i_start + 1, JE), order="F")
    vf = np.zeros((IE, JE - b.j_start + 1), order="F")
    ff = np.zeros((IE, JE), order="F")
    L.orc_itprep(ctypes.byref(g), _d(xx), _d(uf), _d(vf), _d(ff))
    return xx, uf, vf, ff


def troneu(b, coef, b1o=None):
    """Z1O (IE,JE) and B1O after troneu.f90:48-131 with the velocities held by `b`."""
    L = _sor_lib()
    g = _sorgrid(b)
    xx, uf, vf, ff = coef
    b1 = np.zeros((b.IE, b.JE), order="F") if b1o is None else np.array(b1o, order="F", copy=True)
    z = np.zeros((b.IE, b.JE), order="F")
    L.orc_troneu(ctypes.byref(g), _d(xx), _d(uf), _d(vf), _d(ff), _d(b.u1o), _d(b.uzo), _d(b.v1e), _d(b.vze),
                 _d(b1), _d(z))
    return z, b1


# ---- 2-D bottom-pressure solve, CG2D4NONHY (nh_oracle_bot.c) --------------------------------------


def _bot_lib():
    L = lib()
    if not getattr(L, "_bot", False):
        from orctm_amd.abi import GridC

        L.orc_bottom_rhs.argtypes = [ctypes.POINTER(_VGrid)] + [_dp] * 11
        L.orc_op2d_coef.argtypes = [ctypes.POINTER(GridC), _dp]
        L.orc_op2d_apply.argtypes = [ctypes.POINTER(GridC), _dp, _dp, _dp]
        L.orc_solve2d.argtypes = [ctypes.POINTER(GridC), _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_solve2d.restype = ctypes.c_int
        L.orc_mg_apply.argtypes = [ctypes.POINTER(GridC), _dp, _dp, _dp]
        L.orc_mg_apply.restype = ctypes.c_int
        L.orc_bottom_correction.argtypes = [ctypes.POINTER(GridC), _dp, _dp]
        L._bot = True
    return L


def bottom_rhs(grid, v, divg, botgu=None, botgv=None, botgw=None):
    """ocvtot_nh.f90:64-144: (DIVG_bot_surf (IE,JE) on the interior, BOTGU, BOTGV, BOTGW). The BOT* arrays
    are the module arrays the source keeps between steps (zero if not given)."""
    L = _bot_lib()
    g = _vgrid(grid, v)
    IE, JE = v.IE, v.JE
    bu = np.zeros((IE - v.i_start + 1, JE), order="F") if botgu is None else np.array(botgu, order="F", copy=True)
    bv = np.zeros((IE, JE - v.j_start + 1), order="F") if botgv is None else np.array(botgv, order="F", copy=True)
    bw = np.zeros((IE, JE), order="F") if botgw is None else np.array(botgw, order="F", copy=True)
    dbs = np.zeros((IE, JE), order="F")
    dv = np.asfortranarray(divg, dtype=np.float64)
    L.orc_bottom_rhs(ctypes.byref(g), _d(v.uoo), _d(v.uoo_1), _d(v.voe), _d(v.voe_1), _d(v.wno), _d(v.wno_1),
                     _d(dv), _d(bu), _d(bv), _d(bw), _d(dbs))
    return dbs, bu, bv, bw


def op2d_coef(grid):
    """compute_matrix_2d (mo_petsc.f90:401-538, Neumann off): 5 planes (centre, west, east, north, south)
    over the (IE-2) x (JE-2) interior box, DMDA order."""
    L = _bot_lib()
    c = np.zeros((5, grid.nx * grid.ny))
    L.orc_op2d_coef(ctypes.byref(grid.c_struct()), _d(c))
    return c


def op2d_apply(grid, coef, p):
    L = _bot_lib()
    y = np.zeros(grid.nx * grid.ny)
    L.orc_op2d_apply(ctypes.byref(grid.c_struct()), _d(np.ascontiguousarray(coef)), _d(np.ascontiguousarray(p)), _d(y))
    return y


def op2d_scipy(grid, coef):
    import scipy.sparse as sp

    nx, ny = grid.nx, grid.ny
    n = nx * ny
    idx = np.arange(n)
    i, j = idx % nx, idx // nx
    rows, cols, vals = [idx], [idx], [coef[0]]
    for plane, di, dj in ((1, -1, 0), (2, 1, 0), (3, 0, -1), (4, 0, 1)):
        ok = (i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < ny) & (coef[plane] != 0.0)
        rows.append(idx[ok])
        cols.append(idx[ok] + di + dj * nx)
        vals.append(coef[plane][ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def solve2d(grid, coef, b, x0=None, rtol=1e-10, maxit=100000):
    """The device's 2-D PCG (V-cycle preconditioned) restated on the CPU: (x, iterations, rel_res)."""
    L = _bot_lib()
    x = np.zeros(grid.nx * grid.ny) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    rr = ctypes.c_double()
    it = L.orc_solve2d(ctypes.byref(grid.c_struct()), _d(np.ascontiguousarray(coef)), _d(np.ascontiguousarray(b)),
                       _d(x), rtol, maxit, ctypes.byref(rr))
    return x, it, rr.value


def mg_apply(grid, coef, r):
    """One V-cycle of the 2-D solve's multigrid preconditioner: (z, levels)."""
    L = _bot_lib()
    z = np.zeros(grid.nx * grid.ny)
    nl = L.orc_mg_apply(ctypes.byref(grid.c_struct()), _d(np.ascontiguousarray(coef)),
                        _d(np.ascontiguousarray(r, dtype=np.float64)), _d(z))
    return z, nl


def bottom_correction(grid, pnh_bot, divg):
    """ComputeRHS's CG2D4NONHY corrections (mo_petsc.f90:252-374) applied to a copy of DIVG."""
    L = _bot_lib()
    out = np.array(divg, order="F", dtype=np.float64, copy=True)
    L.orc_bottom_correction(ctypes.byref(grid.c_struct()), _d(np.asfortranarray(pnh_bot, dtype=np.float64)), _d(out))
    return out
