"""ctypes mirror of include/orctm_nh.h and the loader of the in-tree HIP library.

The library (``orctm_amd/liborctm_nh.so``) is built by ``orctm_amd/build.py`` with hipcc for
gfx950. Loading fails loudly if it is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ORCTM_NH_LIB", os.path.join(_HERE, "liborctm_nh.so"))

OK, NOT_CONVERGED, BREAKDOWN, ERR_NONFINITE, ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_NODEVICE, PENDING = range(9)
STATUS_NAMES = ["OK", "NOT_CONVERGED", "BREAKDOWN", "ERR_NONFINITE", "ERR_INVALID", "ERR_NOMEM", "ERR_HIP",
                "ERR_NODEVICE", "PENDING"]
FP64, MIXED = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)


class GridC(ctypes.Structure):
    _fields_ = [
        ("ie", ctypes.c_int32), ("je", ctypes.c_int32), ("ke", ctypes.c_int32),
        ("i_start", ctypes.c_int32), ("j_start", ctypes.c_int32),
        ("dlxp", _dp), ("dlyp", _dp), ("dlxu", _dp), ("dlyv", _dp), ("ddpo", _dp), ("weto", _dp), ("dzw", _dp),
        ("cyclic_x", ctypes.c_int32), ("cyclic_y", ctypes.c_int32),
        ("open_w", ctypes.c_int32), ("open_e", ctypes.c_int32), ("open_n", ctypes.c_int32),
        ("open_s", ctypes.c_int32),
    ]


class OptionsC(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32), ("maxit", ctypes.c_int32), ("max_outer", ctypes.c_int32),
        ("warm_start", ctypes.c_int32), ("rtol", ctypes.c_double), ("rtol_inner", ctypes.c_double),
        ("check_every", ctypes.c_int32), ("device", ctypes.c_int32), ("profile", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32), ("bottom_solve", ctypes.c_int32), ("zw_table", ctypes.c_int32),
        ("buffer_addressing", ctypes.c_int32), ("reserved", ctypes.c_int32 * 1),
    ]


class ProfileC(ctypes.Structure):
    _fields_ = [
        ("matvec_ms", ctypes.c_double), ("update_ms", ctypes.c_double), ("matvec_launches", ctypes.c_int64),
        ("update_launches", ctypes.c_int64), ("cells", ctypes.c_int64), ("reserved0", ctypes.c_int64),
        ("reserved", ctypes.c_double * 2),
    ]


class StatsC(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("outer_iters", ctypes.c_int32), ("inner_iters", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32), ("rel_res", ctypes.c_double), ("bnorm", ctypes.c_double),
        ("solve_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("zw_table", ctypes.c_int32),
        ("buffer_addressing", ctypes.c_int32), ("reserved", ctypes.c_double * 3),
    ]

    def as_dict(self):
        return {"status": STATUS_NAMES[self.status] if 0 <= self.status < len(STATUS_NAMES) else self.status,
                "outer_iters": self.outer_iters, "inner_iters": self.inner_iters, "rel_res": self.rel_res,
                "bnorm": self.bnorm, "solve_ms": self.solve_ms, "total_ms": self.total_ms,
                "tile_rows": self.tile_rows, "zw_table": self.zw_table,
                "buffer_addressing": self.buffer_addressing}


class ServerOptionsC(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_active", ctypes.c_int32), ("device", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("max_cohort", ctypes.c_int32), ("cohort_graphs", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 5)]


class RequestStatsC(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("steps_done", ctypes.c_int32), ("inner_iters", ctypes.c_int32),
        ("outer_iters", ctypes.c_int32), ("rel_res", ctypes.c_double), ("t_submit_ms", ctypes.c_double),
        ("t_admit_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double), ("reserved", ctypes.c_double * 4),
    ]

    def as_dict(self):
        return {"status": STATUS_NAMES[self.status] if 0 <= self.status < len(STATUS_NAMES) else self.status,
                "steps_done": self.steps_done, "inner_iters": self.inner_iters, "outer_iters": self.outer_iters,
                "rel_res": self.rel_res, "t_submit_ms": self.t_submit_ms, "t_admit_ms": self.t_admit_ms,
                "t_done_ms": self.t_done_ms}


# Every symbol include/orctm_nh.h declares (checked by the CPU test suite).
class VGridC(ctypes.Structure):
    """orctm_nh_vgrid_t: velocity-side metrics of OCVTOT_NH (Fortran arrays with halos) and DT, DTI."""
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in
                ("dlxp", "dlyp", "ddpo", "weto", "amsuo", "dduo", "dtdxuo", "dlyu", "amsue", "ddue", "dpye",
                 "dlxv", "dtdxpo", "dtdyo")] + [("dt", ctypes.c_double), ("dti", ctypes.c_double)]


class SorGridC(ctypes.Structure):
    """orctm_nh_sor_grid_t: metrics and constants of the barotropic SOR solve (ITPREP / TRONEU)."""
    _fields_ = [(n, ctypes.c_int32) for n in ("ie", "je", "i_start", "j_start", "cyclic_x", "cyclic_y",
                                             "iterations", "device", "half_sweeps", "reserved")] + \
               [(n, ctypes.POINTER(ctypes.c_double)) for n in ("dlxp", "dlyp", "dtdxpo", "dtdyo", "weto", "dlxu",
                                                              "dlyu", "deuto", "dlxv", "dlyv", "deute")] + \
               [(n, ctypes.c_double) for n in ("g", "conn", "cono", "stabn", "dt", "sorpar")]


EXPORTS = [
    "orctm_nh_abi_version", "orctm_nh_device_count", "orctm_nh_default_options", "orctm_nh_status_string",
    "orctm_nh_last_error",
    "orctm_nh_create", "orctm_nh_solve", "orctm_nh_destroy", "orctm_nh_set_rhs", "orctm_nh_set_guess",
    "orctm_nh_solve_resident", "orctm_nh_get_solution", "orctm_nh_apply", "orctm_nh_get_operator",
    "orctm_nh_get_dims", "orctm_nh_device_bytes", "orctm_nh_inner_apply", "orctm_nh_inner_precond",
    "orctm_nh_batch_create", "orctm_nh_batch_solve", "orctm_nh_batch_destroy", "orctm_nh_profile_read",
    "orctm_nh_batch_profile_read", "orctm_nh_clone", "orctm_nh_perturb_rhs", "orctm_nh_copy_probe",
    "orctm_nh_state_bytes", "orctm_nh_server_default_options", "orctm_nh_server_create", "orctm_nh_server_submit",
    "orctm_nh_server_poll", "orctm_nh_server_wait", "orctm_nh_server_drain", "orctm_nh_server_info",
    "orctm_nh_server_destroy", "orctm_nh_attach_vgrid", "orctm_nh_divergence", "orctm_nh_project",
    "orctm_nh_set_profile", "orctm_nh_sor_create", "orctm_nh_troneu", "orctm_nh_sor_info",
    "orctm_nh_sor_destroy", "orctm_nh_get_bottom", "orctm_nh_apply_noalloc", "orctm_nh_bottom_symmetry",
]

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load liborctm_nh.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m orctm_amd.build` "
                           "(hipcc, gfx950). There is no CPU fallback.")
    lib = ctypes.CDLL(path)
    vp, ip = ctypes.c_void_p, ctypes.c_int
    lib.orctm_nh_abi_version.restype = ip
    lib.orctm_nh_device_count.argtypes = [ctypes.POINTER(ctypes.c_int32)]
    lib.orctm_nh_default_options.argtypes = [ctypes.POINTER(OptionsC)]
    lib.orctm_nh_default_options.restype = None
    lib.orctm_nh_status_string.argtypes = [ip]
    lib.orctm_nh_status_string.restype = ctypes.c_char_p
    lib.orctm_nh_last_error.argtypes = []
    lib.orctm_nh_last_error.restype = ctypes.c_char_p
    lib.orctm_nh_create.argtypes = [ctypes.POINTER(GridC), ctypes.POINTER(OptionsC), ctypes.POINTER(vp)]
    lib.orctm_nh_solve.argtypes = [vp, _dp, _dp, ctypes.POINTER(StatsC)]
    lib.orctm_nh_destroy.argtypes = [vp]
    lib.orctm_nh_set_rhs.argtypes = [vp, _dp]
    lib.orctm_nh_set_guess.argtypes = [vp, _dp]
    lib.orctm_nh_solve_resident.argtypes = [vp, ctypes.POINTER(StatsC)]
    lib.orctm_nh_get_solution.argtypes = [vp, _dp]
    lib.orctm_nh_apply.argtypes = [vp, _dp, _dp]
    lib.orctm_nh_apply_noalloc.argtypes = [vp, _dp, _dp]
    lib.orctm_nh_bottom_symmetry.argtypes = [vp, _dp]
    lib.orctm_nh_get_operator.argtypes = [vp, _dp]
    lib.orctm_nh_get_dims.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int32)] * 3
    lib.orctm_nh_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.orctm_nh_inner_apply.argtypes = [vp, vp, vp]
    lib.orctm_nh_inner_precond.argtypes = [vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    lib.orctm_nh_batch_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.orctm_nh_batch_solve.argtypes = [vp, ctypes.POINTER(StatsC)]
    lib.orctm_nh_batch_destroy.argtypes = [vp]
    lib.orctm_nh_clone.argtypes = [vp, ctypes.POINTER(vp)]
    lib.orctm_nh_state_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.orctm_nh_server_default_options.argtypes = [ctypes.POINTER(ServerOptionsC)]
    lib.orctm_nh_server_default_options.restype = None
    lib.orctm_nh_server_create.argtypes = [ctypes.POINTER(ServerOptionsC), ctypes.POINTER(vp)]
    lib.orctm_nh_server_submit.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, _dp,
                                           ctypes.POINTER(ctypes.c_int64)]
    lib.orctm_nh_server_poll.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RequestStatsC)]
    lib.orctm_nh_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RequestStatsC)]
    lib.orctm_nh_server_drain.argtypes = [vp]
    lib.orctm_nh_server_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.orctm_nh_server_destroy.argtypes = [vp]
    lib.orctm_nh_copy_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.orctm_nh_perturb_rhs.argtypes = [vp, ctypes.c_double, ctypes.c_uint64]
    lib.orctm_nh_profile_read.argtypes = [vp, ctypes.POINTER(ProfileC), ctypes.c_int32]
    lib.orctm_nh_batch_profile_read.argtypes = [vp, ctypes.POINTER(ProfileC), ctypes.c_int32]
    for name in EXPORTS:
        f = getattr(lib, name)
        if f.restype is ctypes.c_int and name not in ("orctm_nh_default_options",):
            f.restype = ip
    _lib = lib
    return lib


def copy_probe(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Achievable streaming bandwidth (GB/s) from a plain copy kernel on the device."""
    g = ctypes.c_double()
    rc = load().orctm_nh_copy_probe(device, nbytes, reps, ctypes.byref(g))
    if rc != OK:
        raise RuntimeError(f"orctm_nh_copy_probe: {status_name(rc)} {last_error()}")
    return g.value


def last_error() -> str:
    return load().orctm_nh_last_error().decode(errors="replace")


def status_name(code: int) -> str:
    return STATUS_NAMES[code] if 0 <= code < len(STATUS_NAMES) else f"status {code}"
