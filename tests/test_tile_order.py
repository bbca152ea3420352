"""The order of a streaming launch's tiles (orctm_amd/csrc/nh_tiles.h), compiled for the host.

Block b runs on XCD b % 8. balanced_order must launch every tile exactly once (empty tiles, -1, pad
the list), give each XCD a contiguous run of the natural order (neighbouring tiles share an L2), make
the XCDs' work (levels + 4 per tile) equal to within one tile, put a run's deepest systems first, and
reduce to xcd_order when every system has the same depth or the list is short."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from orctm_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
X = 8


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    so = str(tmp_path_factory.mktemp("tiles") / "libtiles.so")
    subprocess.run([gxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "tiles_host.cpp")],
                   check=True)
    L = ctypes.CDLL(so)
    p = ctypes.POINTER(ctypes.c_int32)
    L.balanced_order.argtypes = [p, ctypes.c_int, p, ctypes.c_int, p, ctypes.c_int]
    L.balanced_order.restype = ctypes.c_int
    L.xcd_order.argtypes = [p, ctypes.c_int, p]
    L.xcd_order.restype = ctypes.c_int
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def balanced(lib, sys, depth, min_tiles=256):
    sys = np.ascontiguousarray(sys, dtype=np.int32)
    depth = np.ascontiguousarray(depth, dtype=np.int32)
    out = np.zeros(4 * len(sys) + 64, dtype=np.int32)
    n = lib.balanced_order(ptr(sys), len(sys), ptr(depth), min_tiles, ptr(out), len(out))
    assert n > 0
    return out[:n]


def xcd(lib, sys):
    sys = np.ascontiguousarray(sys, dtype=np.int32)
    out = np.zeros(len(sys), dtype=np.int32)
    n = lib.xcd_order(ptr(sys), len(sys), ptr(out))
    return out[:n]


def batch_tiles(rows=12):
    """The 64-system batch workload at one tile height: system index per tile (natural order), depths."""
    specs = synth.config4_sizes(n=64)
    sys = []
    for q, sp in enumerate(specs):
        sys += [q] * (((sp.nx + 63) // 64) * ((sp.ny + rows - 1) // rows))
    return np.array(sys), np.array([sp.nz for sp in specs])


def runs(order):
    return [order[x::X] for x in range(X)]


def test_every_tile_once_runs_contiguous_and_work_balanced(lib):
    sys, depth = batch_tiles()
    order = balanced(lib, sys, depth)
    real = order[order >= 0]
    assert len(order) % X == 0
    assert np.array_equal(np.sort(real), np.arange(len(sys)))
    work = lambda idx: float(np.sum(depth[sys[idx]] + 4))
    per = [work(r[r >= 0]) for r in runs(order)]
    biggest = float(np.max(depth) + 4)
    assert max(per) - min(per) <= biggest + 1e-9, per
    # xcd_order on the same tiles: equal tile counts, unequal work
    plain = [work(r) for r in runs(xcd(lib, sys))]
    assert max(plain) / np.mean(plain) > 1.03 and max(per) / np.mean(per) < 1.005
    for r in runs(order):
        r = r[r >= 0]
        # a contiguous run of the natural order, deepest systems first, a system's tiles in order
        assert np.array_equal(np.sort(r), np.arange(r.min(), r.max() + 1))
        d = depth[sys[r]]
        assert np.all(np.diff(d) <= 0)
        for q in np.unique(sys[r]):
            assert np.all(np.diff(r[sys[r] == q]) == 1)
        # empty tiles only after the run's real ones
    for x, r in enumerate(runs(order)):
        k = np.nonzero(r < 0)[0]
        assert len(k) == 0 or np.all(r[k[0]:] < 0), x


def test_equal_depths_give_the_plain_xcd_runs(lib):
    sys = np.zeros(1024, dtype=np.int32)  # the regional tile at 12 rows: 16 x 64 tiles
    order = balanced(lib, sys, [60])
    assert np.array_equal(order, xcd(lib, sys))
    sys = np.repeat(np.arange(5), [300, 200, 250, 141, 129])  # several systems, one depth
    order = balanced(lib, sys, [40] * 5)
    lens = [len(r[r >= 0]) for r in runs(order)]
    assert max(lens) - min(lens) <= 1 and np.count_nonzero(order < 0) < X
    assert np.array_equal(np.concatenate([r[r >= 0] for r in runs(order)]), np.arange(len(sys)))


def test_short_lists_keep_the_plain_order(lib):
    sys = np.repeat(np.arange(3), [40, 100, 60])
    assert np.array_equal(balanced(lib, sys, [30, 60, 45]), xcd(lib, sys))
