"""Host-side mirror of the reference's MO_PETSC interface, over the C ABI of liborctm_nh.so.

Reference (ORCTM/src_orctm/mo_petsc.f90):

* ``petsc_start``   (:41-128)  -> :func:`petsc_start` / :class:`NHSolver` construction
* ``petsc_solve_3D`` (:211-233) -> :func:`petsc_solve_3D` / :meth:`NHSolver.solve`
* ``petsc_end``     (:131-143) -> :func:`petsc_end` / :meth:`NHSolver.close`

Arrays are the caller's Fortran arrays (numpy, order='F', with the halo ring). ``solve`` reads
``DIVG`` and writes the interior of ``PNH`` only, and warm-starts from the previous solution as
``KSPSetInitialGuessNonzero`` does (:232). Unlike the reference (which ignores PETSc's ierr), every
call reports a status; ``solve`` raises :class:`NHError` on errors and returns the statistics.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import abi


class NHError(RuntimeError):
    def __init__(self, code: int, what: str):
        extra = ""
        if code in (abi.ERR_HIP, abi.ERR_NOMEM, abi.ERR_NODEVICE):
            extra = f" [{abi.last_error()}]"
        super().__init__(f"{what}: {abi.status_name(code)} ({code}){extra}")
        self.code = code


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def options(precision: int = abi.MIXED, rtol: float = 1e-10, rtol_inner: Optional[float] = None,
            maxit: int = 10000, max_outer: int = 40, warm_start: bool = True, check_every: int = 16,
            device: int = 0, profile: bool = False, tile_rows: int = 0, bottom_solve: bool = False,
            zw_table: int = 0, buffer_addressing: int = 0) -> abi.OptionsC:
    """Solver options (orctm_nh_options_t). tile_rows: 0 = tuned per size class, or 4, 8, 12.
    bottom_solve: ORCTM's CG2D4NONHY option in divergence() (the 2-D bottom-pressure solve and the
    corrections of DIVG by PNH_bot). zw_table: where the iteration kernels read ZW/ZP, 0 = the
    per-level table when it reproduces every column and the launches cover at least 2^20 cells,
    1 = the table whenever it reproduces every column, -1 = the per-cell planes. buffer_addressing:
    0 = the matvec kernel addresses memory through buffer descriptors where every inner array fits
    2^31 bytes, -1 = always through 64-bit pointers (same bits)."""
    lib = abi.load()
    o = abi.OptionsC()
    lib.orctm_nh_default_options(ctypes.byref(o))
    o.precision = precision
    o.rtol = rtol
    o.rtol_inner = -1.0 if rtol_inner is None else rtol_inner
    o.maxit = maxit
    o.max_outer = max_outer
    o.warm_start = 1 if warm_start else 0
    o.check_every = check_every
    o.device = device
    o.profile = 1 if profile else 0
    o.tile_rows = tile_rows
    o.bottom_solve = 1 if bottom_solve else 0
    o.zw_table = zw_table
    o.buffer_addressing = buffer_addressing
    return o


class NHSolver:
    """One system on one GPU: petsc_start on construction, petsc_solve_3D per time step."""

    def __init__(self, grid, opts: Optional[abi.OptionsC] = None, _handle=None, **kw):
        self.lib = abi.load()
        self.grid = grid
        self.opts = opts if opts is not None else options(**kw)
        if _handle is None:
            h = ctypes.c_void_p()
            rc = self.lib.orctm_nh_create(ctypes.byref(grid.c_struct()), ctypes.byref(self.opts), ctypes.byref(h))
            if rc != abi.OK:
                raise NHError(rc, "orctm_nh_create")
            _handle = h
        self.h = _handle
        self.shape = (grid.IE, grid.JE, grid.KE)
        self.fp32 = self.opts.precision == abi.MIXED

    def clone(self) -> "NHSolver":
        """A solver on the same grid sharing this one's assembled operator (orctm_nh_clone)."""
        h = ctypes.c_void_p()
        self._rc(self.lib.orctm_nh_clone(self.h, ctypes.byref(h)), "orctm_nh_clone")
        return NHSolver(self.grid, self.opts, _handle=h)

    def set_profile(self, on: bool) -> None:
        """Kernel timing on/off between solves (orctm_nh_set_profile)."""
        self._rc(self.lib.orctm_nh_set_profile(self.h, 1 if on else 0), "orctm_nh_set_profile")

    def perturb_rhs(self, amplitude: float, seed: int) -> None:
        self._rc(self.lib.orctm_nh_perturb_rhs(self.h, amplitude, seed), "orctm_nh_perturb_rhs")

    # -- the steps of OCVTOT_NH around the solve (ocvtot_nh.f90:11-62, 151-230) ------------------
    def _vshape(self, kind: str):
        IE, JE, KE = self.shape
        NU, NV = IE - self.grid.i_start + 1, JE - self.grid.j_start + 1
        return {"p2": (IE, JE), "p3": (IE, JE, KE), "u2": (NU, JE), "u3": (NU, JE, KE), "v2": (IE, NV),
                "v3": (IE, NV, KE), "w3": (IE, JE, KE + 1)}[kind]

    def _check_v(self, a: np.ndarray, kind: str, name: str) -> None:
        shp = self._vshape(kind)
        if a.dtype != np.float64 or a.shape != shp or not a.flags["F_CONTIGUOUS"]:
            raise ValueError(f"{name} must be a Fortran-ordered float64 array of shape {shp}")

    _VGRID = (("dlxp", "p2"), ("dlyp", "p2"), ("ddpo", "p3"), ("weto", "p3"), ("amsuo", "u3"), ("dduo", "u3"),
              ("dtdxuo", "u2"), ("dlyu", "u2"), ("amsue", "v3"), ("ddue", "v3"), ("dpye", "v2"), ("dlxv", "v2"),
              ("dtdxpo", "p2"), ("dtdyo", "p2"))

    def attach_vgrid(self, v) -> None:
        """Upload the velocity-side metrics (arrays named as in MO_COMMO1, e.g. from synth.make_velocity)."""
        names = tuple(n for n, _ in self._VGRID)
        for n, kind in self._VGRID:
            self._check_v(getattr(v, n), kind, n)
        vg = abi.VGridC(*[_dp(getattr(v, n)) for n in names], float(v.dt), float(v.dti))
        self._vg_keep = v
        self._rc(self.lib.orctm_nh_attach_vgrid(self.h, ctypes.byref(vg)), "orctm_nh_attach_vgrid")

    def divergence(self, uoo, uoo_1, voe, voe_1, wno, wno_1, zo, divg: Optional[np.ndarray] = None) -> None:
        """DIVG from the velocity tendencies (ocvtot_nh.f90:11-62); the resident b becomes DIVG*WETO. With
        bottom_solve, PNH_bot is solved first and DIVG corrected by it (CG2D4NONHY); `divg` then receives
        the corrected DIVG, as the source's ComputeRHS leaves it."""
        arrs = (uoo, uoo_1, voe, voe_1, wno, wno_1, zo)
        for a, kind, n in zip(arrs, ("u3", "u3", "v3", "v3", "w3", "w3", "p2"),
                              ("uoo", "uoo_1", "voe", "voe_1", "wno", "wno_1", "zo")):
            self._check_v(a, kind, n)
        if divg is not None:
            self._check_f(divg, "divg")
        self._rc(self.lib.orctm_nh_divergence(self.h, *[_dp(a) for a in arrs],
                                              _dp(divg) if divg is not None else None), "orctm_nh_divergence")

    def get_bottom(self) -> dict:
        """With bottom_solve: PNH_bot (IE,JE, halo included), DIVG_bot_surf (IE,JE, interior), and the
        iterations, relative residual and device time of the last 2-D bottom solve (orctm_nh_get_bottom)."""
        IE, JE, KE = self.shape
        p = np.zeros((IE, JE), order="F")
        d = np.zeros((IE, JE), order="F")
        it, rr, ms = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        self._rc(self.lib.orctm_nh_get_bottom(self.h, _dp(p), _dp(d), ctypes.byref(it), ctypes.byref(rr),
                                              ctypes.byref(ms)), "orctm_nh_get_bottom")
        return {"pnh_bot": p, "divg_bot_surf": d, "iterations": it.value, "rel_res": rr.value, "device_ms": ms.value}

    def bottom_symmetry(self) -> float:
        """With bottom_solve: ||K - K^T||_F / ||K||_F of the bottom solve's 2-D operator (orctm_nh_bottom_symmetry)."""
        d = ctypes.c_double()
        self._rc(self.lib.orctm_nh_bottom_symmetry(self.h, ctypes.byref(d)), "orctm_nh_bottom_symmetry")
        return d.value

    def project(self, pnh, uoo, voe, wno) -> None:
        """PNH halo, U/V pressure correction and WNO by continuity (ocvtot_nh.f90:151-230), in place."""
        for a, kind, n in zip((pnh, uoo, voe, wno), ("p3", "u3", "v3", "w3"), ("pnh", "uoo", "voe", "wno")):
            self._check_v(a, kind, n)
        self._rc(self.lib.orctm_nh_project(self.h, _dp(pnh), _dp(uoo), _dp(voe), _dp(wno)), "orctm_nh_project")

    # -- the reference's three calls -------------------------------------------------------------
    def solve(self, divg: np.ndarray, pnh: np.ndarray, check: bool = True) -> dict:
        """petsc_solve_3D: b = DIVG*WETO, solve, write the PNH interior. Returns the statistics."""
        self._check_f(divg, "divg")
        self._check_f(pnh, "pnh")
        st = abi.StatsC()
        rc = self.lib.orctm_nh_solve(self.h, _dp(divg), _dp(pnh), ctypes.byref(st))
        if check and rc not in (abi.OK,):
            raise NHError(rc, "orctm_nh_solve")
        d = st.as_dict()
        d["rc"] = rc
        return d

    def close(self) -> None:
        """petsc_end."""
        if getattr(self, "h", None):
            self.lib.orctm_nh_destroy(self.h)
            self.h = None

    # -- resident (split) form -------------------------------------------------------------------
    def set_rhs(self, divg: np.ndarray) -> None:
        self._check_f(divg, "divg")
        self._rc(self.lib.orctm_nh_set_rhs(self.h, _dp(divg)), "orctm_nh_set_rhs")

    def set_guess(self, pnh: np.ndarray) -> None:
        self._check_f(pnh, "pnh")
        self._rc(self.lib.orctm_nh_set_guess(self.h, _dp(pnh)), "orctm_nh_set_guess")

    def solve_resident(self) -> dict:
        st = abi.StatsC()
        rc = self.lib.orctm_nh_solve_resident(self.h, ctypes.byref(st))
        d = st.as_dict()
        d["rc"] = rc
        return d

    def get_solution(self, pnh: np.ndarray) -> np.ndarray:
        self._check_f(pnh, "pnh")
        self._rc(self.lib.orctm_nh_get_solution(self.h, _dp(pnh)), "orctm_nh_get_solution")
        return pnh

    # -- verification entry points ---------------------------------------------------------------
    def apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._rc(self.lib.orctm_nh_apply(self.h, _dp(x), _dp(y)), "orctm_nh_apply")
        return y

    def apply_noalloc(self, x: np.ndarray) -> np.ndarray:
        """y = A x like apply(), staged in the library's fixed device array (orctm_nh_apply_noalloc)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._rc(self.lib.orctm_nh_apply_noalloc(self.h, _dp(x), _dp(y)), "orctm_nh_apply_noalloc")
        return y

    def operator(self) -> np.ndarray:
        N = self.grid.cells
        A = np.empty((7, N))
        self._rc(self.lib.orctm_nh_get_operator(self.h, _dp(A)), "orctm_nh_get_operator")
        return A

    def inner_apply(self, x: np.ndarray) -> np.ndarray:
        dt = np.float32 if self.fp32 else np.float64
        x = np.ascontiguousarray(x, dtype=dt)
        y = np.empty_like(x)
        self._rc(self.lib.orctm_nh_inner_apply(self.h, x.ctypes.data, y.ctypes.data), "orctm_nh_inner_apply")
        return y

    def inner_precond(self, r: np.ndarray):
        dt = np.float32 if self.fp32 else np.float64
        r = np.ascontiguousarray(r, dtype=dt)
        z = np.empty_like(r)
        rho = ctypes.c_double()
        self._rc(self.lib.orctm_nh_inner_precond(self.h, r.ctypes.data, z.ctypes.data, ctypes.byref(rho)),
                 "orctm_nh_inner_precond")
        return z, rho.value

    def profile(self, reset: bool = True) -> dict:
        p = abi.ProfileC()
        self._rc(self.lib.orctm_nh_profile_read(self.h, ctypes.byref(p), 1 if reset else 0), "orctm_nh_profile_read")
        return {"matvec_ms": p.matvec_ms, "update_ms": p.update_ms, "matvec_launches": p.matvec_launches,
                "update_launches": p.update_launches, "cells": p.cells}

    def device_bytes(self) -> int:
        b = ctypes.c_int64()
        self._rc(self.lib.orctm_nh_device_bytes(self.h, ctypes.byref(b)), "orctm_nh_device_bytes")
        return b.value

    def state_bytes(self) -> int:
        b = ctypes.c_int64()
        self._rc(self.lib.orctm_nh_state_bytes(self.h, ctypes.byref(b)), "orctm_nh_state_bytes")
        return b.value

    # -- helpers ---------------------------------------------------------------------------------
    def _check_f(self, a: np.ndarray, name: str) -> None:
        if a.dtype != np.float64 or a.shape != self.shape or not a.flags["F_CONTIGUOUS"]:
            raise ValueError(f"{name} must be a Fortran-ordered float64 array of shape {self.shape}")

    @staticmethod
    def _rc(rc: int, what: str) -> None:
        if rc != abi.OK:
            raise NHError(rc, what)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NHBatch:
    """Independent systems (ensemble members, nests) advanced together, one launch per phase."""

    def __init__(self, solvers: Sequence[NHSolver]):
        self.lib = abi.load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        b = ctypes.c_void_p()
        rc = self.lib.orctm_nh_batch_create(len(self.solvers), arr, ctypes.byref(b))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_batch_create")
        self.b = b

    def solve(self):
        st = (abi.StatsC * len(self.solvers))()
        rc = self.lib.orctm_nh_batch_solve(self.b, st)
        return rc, [s.as_dict() for s in st]

    def profile(self, reset: bool = True) -> dict:
        p = abi.ProfileC()
        rc = self.lib.orctm_nh_batch_profile_read(self.b, ctypes.byref(p), 1 if reset else 0)
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_batch_profile_read")
        return {"matvec_ms": p.matvec_ms, "update_ms": p.update_ms, "matvec_launches": p.matvec_launches,
                "update_launches": p.update_launches, "cells": p.cells}

    def close(self):
        if getattr(self, "b", None):
            self.lib.orctm_nh_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NHServer:
    """Continuous-batching request loop (orctm_nh_server_*): requests = model solver + time steps,
    admitted by free device memory into cohorts of one size class that run concurrently on lanes."""

    def __init__(self, hbm_budget: int = 0, max_active: int = 64, device: int = 0, lanes: int = 0,
                 max_cohort: int = 0, cohort_graphs: bool = False):
        """lanes: worker lanes running cohorts concurrently (0 = the default, 4); max_cohort: requests
        per cohort at most (0 = the default, 16); cohort_graphs: replay a cohort's iterations from
        captured hipGraphs instead of launching them directly."""
        self.lib = abi.load()
        o = abi.ServerOptionsC()
        self.lib.orctm_nh_server_default_options(ctypes.byref(o))
        o.hbm_budget = hbm_budget
        o.max_active = max_active
        o.device = device
        if lanes:
            o.lanes = lanes
        if max_cohort:
            o.max_cohort = max_cohort
        o.cohort_graphs = 1 if cohort_graphs else 0
        srv = ctypes.c_void_p()
        rc = self.lib.orctm_nh_server_create(ctypes.byref(o), ctypes.byref(srv))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_server_create")
        self.s = srv
        self._keep = {}

    def submit(self, model: NHSolver, nsteps: int, perturb: float = 0.01, seed: int = 0,
               out: Optional[np.ndarray] = None) -> int:
        if out is not None:
            model._check_f(out, "out")
        t = ctypes.c_int64()
        rc = self.lib.orctm_nh_server_submit(self.s, model.h, nsteps, perturb, seed,
                                             _dp(out) if out is not None else None, ctypes.byref(t))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_server_submit")
        self._keep[t.value] = (model, out)
        return t.value

    def poll(self, ticket: int):
        st = abi.RequestStatsC()
        rc = self.lib.orctm_nh_server_poll(self.s, ticket, ctypes.byref(st))
        return rc == abi.OK, st.as_dict()

    def wait(self, ticket: int) -> dict:
        st = abi.RequestStatsC()
        rc = self.lib.orctm_nh_server_wait(self.s, ticket, ctypes.byref(st))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_server_wait")
        return st.as_dict()

    def drain(self) -> None:
        self.lib.orctm_nh_server_drain(self.s)

    def info(self) -> dict:
        r, b = ctypes.c_int64(), ctypes.c_int64()
        a, q = ctypes.c_int32(), ctypes.c_int32()
        self.lib.orctm_nh_server_info(self.s, ctypes.byref(r), ctypes.byref(b), ctypes.byref(a), ctypes.byref(q))
        return {"rounds": r.value, "bytes_in_use": b.value, "active": a.value, "queued": q.value}

    def close(self) -> None:
        if getattr(self, "s", None):
            self.lib.orctm_nh_server_destroy(self.s)
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Module-level mirror of MO_PETSC for callers written against the reference's subroutines.
_state = {}


def petsc_start(grid, **kw) -> NHSolver:
    """mo_petsc.f90:41-128: create the solver and assemble the operator once."""
    s = NHSolver(grid, **kw)
    _state["solver"] = s
    return s


def petsc_solve_3D(divg: np.ndarray, pnh: np.ndarray) -> dict:
    """mo_petsc.f90:211-233: solve for PNH from DIVG, warm-started from the previous solution."""
    return _state["solver"].solve(divg, pnh)


def petsc_end() -> None:
    """mo_petsc.f90:131-143."""
    s = _state.pop("solver", None)
    if s is not None:
        s.close()


class NHBarotropic:
    """The barotropic free-surface solve (ITPREP at construction, TRONEU per call) on one GPU.
    `b` holds the Fortran arrays with halos and the constants (e.g. from synth.make_barotropic)."""

    _PTRS = ("dlxp", "dlyp", "dtdxpo", "dtdyo", "weto", "dlxu", "dlyu", "deuto", "dlxv", "dlyv", "deute")

    def __init__(self, b, weto=None, device: int = 0, half_sweeps: int = 0):
        self.lib = abi.load()
        self.b = b
        arrs = {n: getattr(b, n) for n in self._PTRS if n != "weto"}
        arrs["weto"] = b.weto1 if weto is None else weto
        for n, a in arrs.items():
            if a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
                raise ValueError(f"{n} must be a Fortran-ordered float64 array")
        self._keep = arrs
        g = abi.SorGridC(b.IE, b.JE, b.i_start, b.j_start, int(b.cyc_x), int(b.cyc_y), int(b.iterations), device,
                         int(half_sweeps), 0, *[_dp(arrs[n]) for n in self._PTRS], float(b.g), float(b.conn), float(b.cono),
                         float(b.stabn), float(b.dt), float(b.sorpar))
        h = ctypes.c_void_p()
        rc = self.lib.orctm_nh_sor_create(ctypes.byref(g), ctypes.byref(h))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_sor_create")
        self.h = h

    def troneu(self, u1o, uzo, v1e, vze, b1o, z1o) -> float:
        """Z1O (and B1O) of troneu.f90:48-131, in place; returns the device time in ms."""
        for a in (u1o, uzo, v1e, vze, b1o, z1o):
            if a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
                raise ValueError("arrays must be Fortran-ordered float64")
        ms = ctypes.c_double()
        rc = self.lib.orctm_nh_troneu(self.h, _dp(u1o), _dp(uzo), _dp(v1e), _dp(vze), _dp(b1o), _dp(z1o),
                                      ctypes.byref(ms))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_troneu")
        return ms.value

    def info(self) -> dict:
        """Half-sweeps per blocked launch in use, and kernel launches per call's sweeps."""
        hs, nl = ctypes.c_int32(), ctypes.c_int32()
        rc = self.lib.orctm_nh_sor_info(self.h, ctypes.byref(hs), ctypes.byref(nl))
        if rc != abi.OK:
            raise NHError(rc, "orctm_nh_sor_info")
        return {"half_sweeps": hs.value, "launches": nl.value}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.orctm_nh_sor_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

