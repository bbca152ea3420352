"""The C-ABI library loads on a machine without a GPU and exports every symbol include/orctm_nh.h
declares; entry points that need no device behave as documented."""
import ctypes
import os
import re

import pytest

from orctm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "orctm_nh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(orctm_nh_\w+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert declared_symbols() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("liborctm_nh.so not built (python -m orctm_amd.build)")
    lib = abi.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = os.popen(f"nm -D --defined-only {abi.LIB_PATH}").read()
    for name in declared_symbols():
        assert re.search(rf"\bT {name}\b", out), name


def test_host_only_entry_points():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("liborctm_nh.so not built")
    lib = abi.load()
    assert lib.orctm_nh_abi_version() == 1
    o = abi.OptionsC()
    lib.orctm_nh_default_options(ctypes.byref(o))
    assert o.precision == abi.MIXED and o.rtol == 1e-10 and o.maxit == 10000 and o.warm_start == 1
    assert lib.orctm_nh_status_string(0) == b"converged"
    # invalid arguments are rejected before any device call
    h = ctypes.c_void_p()
    assert lib.orctm_nh_create(None, None, ctypes.byref(h)) == abi.ERR_INVALID
    assert lib.orctm_nh_solve(None, None, None, None) == abi.ERR_INVALID
    assert lib.orctm_nh_destroy(None) == abi.ERR_INVALID


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """ctypes mirrors of the ABI structs have the sizes and offsets gcc gives the header's structs."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    checks = {"orctm_nh_grid_t": (abi.GridC, ["ie", "dlxp", "dzw", "cyclic_x", "open_s"]),
              "orctm_nh_options_t": (abi.OptionsC, ["precision", "rtol", "check_every", "profile", "tile_rows",
                                                        "bottom_solve", "zw_table"]),
              "orctm_nh_stats_t": (abi.StatsC, ["status", "tile_rows", "rel_res", "total_ms", "zw_table"]),
              "orctm_nh_profile_t": (abi.ProfileC, ["matvec_ms", "cells"]),
              "orctm_nh_server_options_t": (abi.ServerOptionsC, ["hbm_budget", "max_active", "device", "lanes",
                                                                    "max_cohort", "cohort_graphs"]),
              "orctm_nh_request_stats_t": (abi.RequestStatsC, ["status", "rel_res", "t_done_ms"]),
              "orctm_nh_vgrid_t": (abi.VGridC, ["dlxp", "amsuo", "dtdyo", "dt", "dti"]),
              "orctm_nh_sor_grid_t": (abi.SorGridC, ["iterations", "device", "half_sweeps", "dlxp", "deute", "g",
                                                     "sorpar"])}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/orctm_nh.h"', "int main(void){"]
    for cname, (_, fields) in checks.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for cname, (cls, fields) in checks.items():
        want.append(ctypes.sizeof(cls))
        want += [getattr(cls, f).offset for f in fields]
    assert got == want
