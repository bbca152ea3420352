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
    L.pair_tiles.argtypes = [p, p, ctypes.c_int, p, ctypes.c_int]
    L.pair_tiles.restype = ctypes.c_int
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


def pairs(lib, tx, ty):
    tx = np.asarray(tx, dtype=np.int32)
    ty = np.asarray(ty, dtype=np.int32)
    cap = int((tx * ty).sum())
    out = np.zeros(4 * cap, dtype=np.int32)
    n = lib.pair_tiles(ptr(tx), ptr(ty), len(tx), ptr(out), cap)
    assert n >= 0
    return out[:4 * n].reshape(n, 4)


@pytest.mark.parametrize("tx,ty", [(1, 1), (1, 2), (1, 7), (2, 5), (3, 1), (3, 2), (3, 9), (5, 4), (8, 3), (7, 11)])
def test_pair_blocks_cover_every_tile_once(lib, tx, ty):
    """Each block covers its tile and, horizontally, (bx+1, by) where that exists, or, vertically,
    (bx, by+1); every tile of the system is covered exactly once; vertical blocks only pair the last
    tiles of odd rows, from an even row-block down to the next."""
    got = pairs(lib, [tx], [ty])
    seen = np.zeros((ty, tx), dtype=int)
    for q, bx, by, vert in got:
        assert q == 0 and bx % 2 == 0
        seen[by, bx] += 1
        if vert:
            assert tx % 2 == 1 and bx == tx - 1 and by % 2 == 0 and by + 1 < ty
            seen[by + 1, bx] += 1
        elif bx + 1 < tx:
            seen[by, bx + 1] += 1
    assert (seen == 1).all(), seen
    half_filled = sum(1 for q, bx, by, vert in got if not vert and bx + 1 == tx)
    assert half_filled == (ty % 2 if tx % 2 else 0)  # only the last row-block's tile of an odd row is alone


def test_pair_blocks_of_the_64_system_batch(lib):
    """The batch workload: per system in turn, every tile once; odd rows pair their last tiles."""
    sizes = synth.config4_sizes()
    tx = [-(-s.nx // 64) for s in sizes]
    ty = [-(-s.ny // 12) for s in sizes]
    got = pairs(lib, tx, ty)
    assert (np.diff(got[:, 0]) >= 0).all()  # system-major, as the tiles came
    for q in range(len(sizes)):
        mine = got[got[:, 0] == q]
        covered = sum(1 + (1 if v else (1 if bx + 1 < tx[q] else 0)) for _, bx, _, v in mine)
        assert covered == tx[q] * ty[q]
        assert len(mine) == ty[q] * (tx[q] // 2) + ((ty[q] + 1) // 2 if tx[q] % 2 else 0)
