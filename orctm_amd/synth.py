"""Seeded synthetic inputs for the nonhydrostatic pressure Poisson solve.

Produces exactly the arrays ORCTM hands to its solver, in the caller's Fortran layout
(column-major, one-cell halo ring, ORCTM/src_orctm/mo_commo1.f90:105-146):

* vertical grid ``DZW(KE)`` stretched with a tanh profile; ``TIESTW``/``TIESTU`` as
  ORCTM/src_orctm/beleg.f90:87-105;
* land mask ``WETO`` and cell thickness ``DDPO`` by the depth rule of
  ORCTM/src_orctm/boden.f90:182-190 (``WETO=1`` iff ``TIESTU(K) < DEPTO``), except that the bottom
  wet cell gets a seeded fraction ``f*DZW`` (the partial-cell "face fraction" of the workloads);
* metrics ``DLXP, DLYP, DLXU(I_start:IE,JE), DLYV(IE,J_start:JE)``;
* ``DIVG``, the divergence of a seeded random velocity field, by the formula of
  ORCTM/src_orctm/ocvtot_nh.f90:30-62 (with DT = 1, ZO = 0).

Random numbers come from a counter-based SplitMix64 stream (vectorised in numpy) and
Box--Muller normals, so every platform produces the same bits for the same seed.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Optional

import numpy as np

BASE_SEED = 20261202

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        k = np.array([(seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
        return _mix(k + _G)[0]


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniforms in [0, 1) from SplitMix64 counters offset..offset+n-1 of (seed, stream)."""
    key = _key(seed, stream)
    with np.errstate(over="ignore"):
        c = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = _mix(key + c * _G)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normals (Box--Muller on two uniform streams)."""
    u1 = 1.0 - uniform(seed, stream, n)  # (0, 1]
    u2 = uniform(seed, stream + 7919, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def stretched_dzw(ke: int, dz0: float, dz1: float, depth: float) -> np.ndarray:
    """DZW(k) = dz0 + (dz1-dz0)/2 (1 + tanh((k-k0)/w)), w = ke/8, k0 by bisection so sum = depth."""
    k = np.arange(ke, dtype=np.float64)
    w = max(ke / 8.0, 1e-3)

    def prof(k0):
        return dz0 + (dz1 - dz0) * 0.5 * (1.0 + np.tanh((k - k0) / w))

    lo, hi = -4.0 * ke - 10, 5.0 * ke + 10  # sum decreases with k0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if prof(mid).sum() > depth:
            lo = mid
        else:
            hi = mid
    return prof(0.5 * (lo + hi))


@dataclasses.dataclass
class GridSpec:
    """Interior sizes nx, ny, nz (the DMDA of mo_petsc.f90:48-50) and the recipe of one system."""

    nx: int
    ny: int
    nz: int
    dx: float = 1850.0
    dy: float = 1850.0
    dz0: float = 5.0
    dz1: float = 250.0
    depth: float = 5000.0
    bathy: str = "random"  # "random" | "ridge" | "flat"
    ridge_height: float = 800.0
    ridge_width: float = 5000.0
    land_frac: float = 0.25
    dmin: float = 50.0
    dmax: float = 5000.0
    smooth: float = 16.0
    cyclic_x: bool = False
    cyclic_y: bool = False
    open_w: bool = True
    open_e: bool = True
    open_n: bool = True
    open_s: bool = True
    frac_lo: float = 0.1
    sigma_u: float = 0.2
    seed: int = BASE_SEED
    dx_profile: Optional[np.ndarray] = None  # optional DLXP(i) along x for the interior, length nx
    lat_deg: Optional[tuple] = None  # (lat0, lat1): a spherical grid, DLXP scaled by cos(latitude), the
    #                                  latitude linear in j from lat0 at J=1 to lat1 at J=JE

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def config1(seed: int = BASE_SEED) -> GridSpec:
    """ISW transect 512 x 4 (periodic y) x 100, dx = 50 m, dz 2 -> 40 m over 1500 m, Gaussian ridge."""
    return GridSpec(512, 4, 100, dx=50.0, dy=50.0, dz0=2.0, dz1=40.0, depth=1500.0, bathy="ridge",
                    ridge_height=800.0, ridge_width=5000.0, cyclic_y=True, open_w=False, open_e=False,
                    open_n=False, open_s=False, frac_lo=0.2, sigma_u=0.1, seed=seed)


def config2(seed: int = BASE_SEED) -> GridSpec:
    """Regional tile 1024 x 768 x 60, dx = dy = 1.85 km, dz 5 -> 250 m over 5000 m, 25 % land."""
    return GridSpec(1024, 768, 60, seed=seed)


def config4_sizes(seed: int = BASE_SEED, n: int = 64):
    """Batch of n systems: nx, ny ~ U{96..512}, nz ~ U{30..60}, seeds seed+1 .. seed+n."""
    u = uniform(seed, 404, 3 * n).reshape(n, 3)
    out = []
    for s in range(n):
        nx = 96 + int(u[s, 0] * (512 - 96 + 1))
        ny = 96 + int(u[s, 1] * (512 - 96 + 1))
        nz = 30 + int(u[s, 2] * (60 - 30 + 1))
        out.append(GridSpec(nx, ny, nz, seed=seed + 1 + s))
    return out


def config5_size(t: float):
    """Log-uniform serving size: nx = 128*16^t, ny = 128*12^t, nz = 30*(8/3)^t."""
    return (int(round(128 * 16.0 ** t)), int(round(128 * 12.0 ** t)), int(round(30 * (8.0 / 3.0) ** t)))


class Grid:
    """Fortran-layout arrays of one system (numpy, order='F'), with a ctypes view for the C ABI."""

    def __init__(self, spec: GridSpec, dzw, dlxp, dlyp, dlxu, dlyv, ddpo, weto):
        self.spec = spec
        self.IE, self.JE, self.KE = spec.nx + 2, spec.ny + 2, spec.nz
        self.i_start = 0
        self.j_start = 0
        self.dzw = np.ascontiguousarray(dzw, dtype=np.float64)
        self.dlxp = np.asfortranarray(dlxp, dtype=np.float64)
        self.dlyp = np.asfortranarray(dlyp, dtype=np.float64)
        self.dlxu = np.asfortranarray(dlxu, dtype=np.float64)
        self.dlyv = np.asfortranarray(dlyv, dtype=np.float64)
        self.ddpo = np.asfortranarray(ddpo, dtype=np.float64)
        self.weto = np.asfortranarray(weto, dtype=np.float64)
        self._c = None

    @property
    def nx(self):
        return self.spec.nx

    @property
    def ny(self):
        return self.spec.ny

    @property
    def nz(self):
        return self.spec.nz

    @property
    def cells(self):
        return self.spec.cells

    def interior(self, a: np.ndarray) -> np.ndarray:
        """Interior of a Fortran (IE,JE,KE) array as a DMDA natural-order vector (i fastest)."""
        return np.ascontiguousarray(a[1:-1, 1:-1, :].transpose(2, 1, 0)).reshape(-1)

    def to_fortran(self, x: np.ndarray, out: Optional[np.ndarray] = None) -> np.ndarray:
        """Natural-order interior vector -> (IE,JE,KE) Fortran array (halo zero unless out is given)."""
        if out is None:
            out = np.zeros((self.IE, self.JE, self.KE), dtype=np.float64, order="F")
        out[1:-1, 1:-1, :] = x.reshape(self.nz, self.ny, self.nx).transpose(2, 1, 0)
        return out

    def c_struct(self):
        from .abi import GridC

        if self._c is None:
            d = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            s = self.spec
            self._c = GridC(self.IE, self.JE, self.KE, self.i_start, self.j_start, d(self.dlxp), d(self.dlyp),
                            d(self.dlxu), d(self.dlyv), d(self.ddpo), d(self.weto), d(self.dzw), int(s.cyclic_x),
                            int(s.cyclic_y), int(s.open_w), int(s.open_e), int(s.open_n), int(s.open_s))
        return self._c


def _halo_2d(a: np.ndarray, spec: GridSpec, closed_value: float) -> None:
    """Fill the halo ring of a (IE,JE) field: cyclic copy, open = copy of the boundary column, else closed."""
    if spec.cyclic_x:
        a[0, :] = a[-2, :]
        a[-1, :] = a[1, :]
    else:
        a[0, :] = a[1, :] if spec.open_w else closed_value
        a[-1, :] = a[-2, :] if spec.open_e else closed_value
    if spec.cyclic_y:
        a[:, 0] = a[:, -2]
        a[:, -1] = a[:, 1]
    else:
        a[:, 0] = a[:, 1] if spec.open_n else closed_value
        a[:, -1] = a[:, -2] if spec.open_s else closed_value


def _gaussian_smooth_periodic(f: np.ndarray, sigma: float) -> np.ndarray:
    if sigma <= 0:
        return f
    nx, ny = f.shape
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.rfftfreq(ny)[None, :]
    g = np.exp(-2.0 * (math.pi * sigma) ** 2 * (kx * kx + ky * ky))
    return np.fft.irfft2(np.fft.rfft2(f) * g, s=f.shape)


def make_grid(spec: GridSpec) -> Grid:
    IE, JE, KE = spec.nx + 2, spec.ny + 2, spec.nz
    dzw = stretched_dzw(KE, spec.dz0, spec.dz1, spec.depth)
    # beleg.f90:87-105
    tiestw = np.zeros(KE + 1)
    tiestw[1:] = np.cumsum(dzw)
    tiestu = np.empty(KE + 1)
    tiestu[:KE] = 0.5 * (tiestw[:KE] + tiestw[1:])
    tiestu[KE] = 9000.0

    # metrics (uniform, or a prescribed DLXP(i) profile)
    if spec.dx_profile is not None:
        dxi = np.empty(IE)
        dxi[1:-1] = np.asarray(spec.dx_profile, dtype=np.float64)
        dxi[0], dxi[-1] = (dxi[-2], dxi[1]) if spec.cyclic_x else (dxi[1], dxi[-2])
    else:
        dxi = np.full(IE, spec.dx)
    dlxp = np.asfortranarray(np.repeat(dxi[:, None], JE, axis=1))
    if spec.lat_deg is not None:
        lat = np.radians(np.linspace(spec.lat_deg[0], spec.lat_deg[1], JE))
        dlxp = np.asfortranarray(dlxp * np.cos(lat)[None, :])
    dlyp = np.full((IE, JE), spec.dy, order="F")
    dlxu = np.empty((IE + 1, JE), order="F")  # I = 0..IE
    dlxu[1:IE, :] = 0.5 * (dlxp[:-1, :] + dlxp[1:, :])
    dlxu[0, :] = dlxp[0, :]
    dlxu[IE, :] = dlxp[IE - 1, :]
    dlyv = np.full((IE, JE + 1), spec.dy, order="F")

    # bathymetry DEPTO(IE,JE)
    if spec.bathy == "ridge":
        x = (np.arange(IE) - 0.5) * spec.dx
        xc = 0.5 * spec.nx * spec.dx
        prof = spec.depth - spec.ridge_height * np.exp(-(((x - xc) / spec.ridge_width) ** 2))
        depto = np.asfortranarray(np.repeat(prof[:, None], JE, axis=1))
    elif spec.bathy == "flat":
        depto = np.full((IE, JE), spec.depth, order="F")
    else:
        noise = normal(spec.seed, 11, spec.nx * spec.ny).reshape(spec.ny, spec.nx).T
        sm = _gaussian_smooth_periodic(noise, spec.smooth)
        ranks = np.empty(sm.size)
        ranks[np.argsort(sm, axis=None, kind="stable")] = np.arange(sm.size)
        q = spec.land_frac * sm.size
        d = np.where(ranks < q, 0.0,
                     spec.dmin + (spec.dmax - spec.dmin) * (ranks - q) / max(sm.size - 1 - q, 1.0))
        depto = np.zeros((IE, JE), order="F")
        depto[1:-1, 1:-1] = d.reshape(spec.nx, spec.ny)
    _halo_2d(depto, spec, 0.0)

    # boden.f90:182-190 mask and thickness; bottom wet cell gets f*DZW, f ~ U(frac_lo, 1]
    kb = np.searchsorted(tiestu[:KE], depto, side="left")  # number of levels with TIESTU(K) < DEPTO
    kb = np.minimum(kb, KE)
    uf = uniform(spec.seed, 23, spec.nx * spec.ny).reshape(spec.ny, spec.nx).T
    f = np.ones((IE, JE), order="F")
    f[1:-1, 1:-1] = 1.0 - uf * (1.0 - spec.frac_lo)
    _halo_2d(f, spec, 1.0)
    lev = np.arange(KE)[None, None, :]
    kb3 = kb[:, :, None]
    weto = (lev < kb3).astype(np.float64)
    ddpo = np.where(lev < kb3 - 1, dzw[None, None, :], 0.0)
    ddpo = np.where(lev == kb3 - 1, f[:, :, None] * dzw[None, None, :], ddpo)
    weto = np.asfortranarray(weto)
    ddpo = np.asfortranarray(ddpo)
    g = Grid(spec, dzw, dlxp, dlyp, dlxu, dlyv, ddpo, weto)
    g.depto, g.tiestw, g.tiestu = depto, tiestw, tiestu
    return g


def make_divg(grid: Grid, seed: Optional[int] = None, sigma: Optional[float] = None) -> np.ndarray:
    """DIVG(IE,JE,KE): divergence of a seeded N(0, sigma) velocity (ocvtot_nh.f90:30-62, DT=1, ZO=0)."""
    spec = grid.spec
    seed = spec.seed if seed is None else seed
    sigma = spec.sigma_u if sigma is None else sigma
    IE, JE, KE = grid.IE, grid.JE, grid.KE
    w3 = grid.weto
    # u at east faces, UOO(I_start:IE,JE,KE), I = 0..IE; masked by both adjacent p-cells
    gu = np.zeros((IE + 1, JE, KE), order="F")
    gu[1:IE, :, :] = sigma * normal(seed, 31, (IE - 1) * JE * KE).reshape(KE, JE, IE - 1).transpose(2, 1, 0)
    gu[1:IE, :, :] *= w3[:-1, :, :] * w3[1:, :, :]
    gu[IE, :, :] = 0.0
    # v at south faces, VOE(IE,J_start:JE,KE), J = 0..JE
    gv = np.zeros((IE, JE + 1, KE), order="F")
    gv[:, 1:JE, :] = sigma * normal(seed, 37, IE * (JE - 1) * KE).reshape(KE, JE - 1, IE).transpose(2, 1, 0)
    gv[:, 1:JE, :] *= w3[:, :-1, :] * w3[:, 1:, :]
    # w at the top of each cell, WNO(IE,JE,KEP); GW(KEP) = 0
    gw = np.zeros((IE, JE, KE + 1), order="F")
    gw[:, :, :KE] = sigma * normal(seed, 41, IE * JE * KE).reshape(KE, JE, IE).transpose(2, 1, 0)
    gw[:, :, 0] *= w3[:, :, 0]
    gw[:, :, 1:KE] *= w3[:, :, :-1] * w3[:, :, 1:]
    divgu = (gu[1:, :, :] - gu[:-1, :, :]) / grid.dlxp[:, :, None]          # I = 1..IE
    divgv = (gv[:, :-1, :] - gv[:, 1:, :]) / grid.dlyp[:, :, None]          # J = 1..JE
    divgw = (gw[:, :, :KE] - gw[:, :, 1:]) / np.maximum(grid.ddpo, 1.0e-3)
    return np.asfortranarray(divgu + divgv + divgw)


def _exch(a: np.ndarray, spec: GridSpec, i0: int = 1, j0: int = 1) -> None:
    """bounds_exch (mo_parallel.f90:551-668) for one process, in place, on an array a(i0:IE, j0:JE[, k])
    (numpy index = Fortran index - lower bound): with ICYCLI_X, A(1) = A(IE-1) and A(IE) = A(2); then
    with ICYCLI_Y, A(j0) = A(JE-1) and A(JE) = A(j0+1) (j0 = 0 for v-points: A(0) and A(1) are used)."""
    IE = spec.nx + 2
    if spec.cyclic_x:
        xs1, xs2 = a[2 - i0].copy(), a[IE - 1 - i0].copy()
        a[1 - i0] = xs2
        a[IE - i0] = xs1
    if spec.cyclic_y:
        y1, y2 = a[:, 1].copy(), a[:, -2].copy()
        a[:, 0] = y2
        a[:, -1] = y1


class VelocityState:
    """Velocity-side metrics and a seeded velocity state of OCVTOT_NH (ocvtot_nh.f90:11-62, 151-236),
    all Fortran arrays with halos (mo_commo1.f90:105-146), I_start = J_start = 0:
    u-points (0:IE, JE[, KE]), v-points (IE, 0:JE[, KE]), w-points (IE, JE, KEP)."""


def make_velocity(grid: Grid, seed: Optional[int] = None, dt: float = 60.0, sigma: Optional[float] = None):
    """Metrics by the model's own rules -- DEUTO = min(DEPTO(I), DEPTO(I+1)), DEUTE likewise in j, end
    columns copied (boden.f90:116-133); DDUO, DDUE, AMSUO, AMSUE from them (boden.f90:195-211);
    DTDXPO = DT/DLXP, DTDYO = DT/DLYP, DTDXUO = DT/DLXU, DPYE = DT/DLYV (beleg.f90:48-74); DTI = 1/DT --
    and seeded velocities UOO, VOE, WNO (masked by AMSUO, AMSUE and wet tops) with their previous-step
    values UOO_1, VOE_1, WNO_1 and a small sea-surface elevation ZO."""
    spec = grid.spec
    seed = spec.seed if seed is None else seed
    sigma = spec.sigma_u if sigma is None else sigma
    IE, JE, KE = grid.IE, grid.JE, grid.KE
    dzw, tiestw, tiestu, depto = grid.dzw, grid.tiestw, grid.tiestu, grid.depto
    v = VelocityState()
    v.IE, v.JE, v.KE, v.i_start, v.j_start = IE, JE, KE, 0, 0
    v.dt, v.dti = float(dt), 1.0 / float(dt)
    deuto = np.zeros((IE + 1, JE), order="F")           # I = 0..IE
    deuto[1:IE, :] = np.minimum(depto[:-1, :], depto[1:, :])
    deuto[0, :] = deuto[1, :]
    deuto[IE, :] = deuto[IE - 1, :]
    _exch(deuto, spec, i0=0)
    deute = np.zeros((IE, JE + 1), order="F")           # J = 0..JE
    deute[:, 1:JE] = np.minimum(depto[:, :-1], depto[:, 1:])
    deute[:, 0] = deute[:, 1]
    deute[:, JE] = deute[:, JE - 1]
    _exch(deute, spec, j0=0)

    def layers(deut):
        d = np.zeros(deut.shape + (KE,), order="F")
        for k in range(KE):
            full = tiestu[k + 1] < deut
            part = (tiestu[k] < deut) & ~full
            d[..., k] = np.where(full, dzw[k], np.where(part, deut - tiestw[k], 0.0))
        return np.asfortranarray(d)

    v.deuto, v.deute = deuto, deute
    v.dduo, v.ddue = layers(deuto), layers(deute)
    v.amsuo = np.asfortranarray((v.dduo > 0.0).astype(np.float64))
    v.amsue = np.asfortranarray((v.ddue > 0.0).astype(np.float64))
    v.dlxp, v.dlyp, v.ddpo, v.weto = grid.dlxp, grid.dlyp, grid.ddpo, grid.weto
    v.dlyu = np.full((IE + 1, JE), spec.dy, order="F")
    v.dlxv = np.empty((IE, JE + 1), order="F")
    v.dlxv[:, 1:] = grid.dlxp
    v.dlxv[:, 0] = grid.dlxp[:, 0]
    v.dtdxpo = np.asfortranarray(dt / grid.dlxp)
    v.dtdyo = np.asfortranarray(dt / grid.dlyp)
    v.dtdxuo = np.asfortranarray(dt / grid.dlxu)
    v.dpye = np.asfortranarray(dt / grid.dlyv)

    def field(stream, shape, mask, scale):
        n = int(np.prod(shape))
        a = scale * normal(seed, stream, n).reshape(shape[::-1]).transpose(tuple(range(len(shape)))[::-1])
        return np.asfortranarray(a * mask)

    wtop = np.zeros((IE, JE, KE + 1), order="F")
    wtop[:, :, :KE] = grid.weto
    v.uoo = field(61, (IE + 1, JE, KE), v.amsuo, sigma)
    v.uoo_1 = field(62, (IE + 1, JE, KE), v.amsuo, sigma)
    v.voe = field(63, (IE, JE + 1, KE), v.amsue, sigma)
    v.voe_1 = field(64, (IE, JE + 1, KE), v.amsue, sigma)
    v.wno = field(65, (IE, JE, KE + 1), wtop, 0.01 * sigma)
    v.wno_1 = field(66, (IE, JE, KE + 1), wtop, 0.01 * sigma)
    v.zo = field(67, (IE, JE), grid.weto[:, :, 0], 0.5)
    for a in (v.uoo, v.uoo_1):
        _exch(a, spec, i0=0)
    for a in (v.voe, v.voe_1):
        _exch(a, spec, j0=0)
    for a in (v.wno, v.wno_1, v.zo):
        _exch(a, spec)
    return v


class BarotropicState:
    """Inputs of the barotropic free-surface solve (itprep.f90, troneu.f90) as Fortran arrays with
    halos, I_start = J_start = 0: metrics, constants and a seeded velocity state."""


def make_barotropic(grid: Grid, seed: Optional[int] = None, dt: float = 60.0, sorpar: float = 1.75,
                    iterations: int = 250):
    """Metrics by the model's rules (u/v-point depths DEUTO, DEUTE as in boden.f90:116-133; DTDXPO,
    DTDYO from beleg.f90:48-49; G = 9.81, CONN = 0.5, CONO = 1 - CONN, STABN = 0.6 as orctm.f90:657-661)
    and seeded barotropic velocities U1O, UZO (masked by AMSUO(:,:,1)) and V1E, VZE (by AMSUE(:,:,1))."""
    spec = grid.spec
    seed = spec.seed if seed is None else seed
    v = make_velocity(grid, seed=seed, dt=dt)
    IE, JE = grid.IE, grid.JE
    b = BarotropicState()
    b.IE, b.JE, b.i_start, b.j_start = IE, JE, 0, 0
    b.cyc_x, b.cyc_y, b.iterations = bool(spec.cyclic_x), bool(spec.cyclic_y), int(iterations)
    b.g, b.conn, b.stabn, b.dt, b.sorpar = 9.81, 0.5, 0.6, float(dt), float(sorpar)
    b.cono = 1.0 - b.conn
    b.dlxp, b.dlyp, b.dtdxpo, b.dtdyo = grid.dlxp, grid.dlyp, v.dtdxpo, v.dtdyo
    b.weto1 = np.asfortranarray(grid.weto[:, :, 0])
    b.dlxu, b.dlyu, b.deuto = grid.dlxu, v.dlyu, v.deuto
    b.dlyv, b.dlxv, b.deute = grid.dlyv, v.dlxv, v.deute

    def field(stream, shape, mask, scale):
        n = int(np.prod(shape))
        a = scale * normal(seed, stream, n).reshape(shape[::-1]).T
        return np.asfortranarray(a * mask)

    # transports exist where the face has a wet top layer (AMSUO(:,:,1), AMSUE(:,:,1), boden.f90:195-211)
    mu = v.amsuo[:, :, 0]
    mv = v.amsue[:, :, 0]
    b.u1o = field(81, (IE + 1, JE), mu, spec.sigma_u)
    b.uzo = field(82, (IE + 1, JE), mu, spec.sigma_u)
    b.v1e = field(83, (IE, JE + 1), mv, spec.sigma_u)
    b.vze = field(84, (IE, JE + 1), mv, spec.sigma_u)
    for a in (b.u1o, b.uzo):
        _exch(a, spec, i0=0)
    for a in (b.v1e, b.vze):
        _exch(a, spec, j0=0)
    return b


def build(spec: GridSpec, with_rhs: bool = True):
    g = make_grid(spec)
    return (g, make_divg(g)) if with_rhs else g
