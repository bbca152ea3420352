"""orctm_nh_apply_noalloc, the variant of orctm_nh_apply that allocates no device memory (its x and y
are staged in a fixed device array of the library): through the C ABI it gives the same bits as
orctm_nh_apply and the oracle's FP64 matvec, on two calls in a row (the second on a smaller system,
so stale staging from the first would show), and it refuses a system beyond its staging limit before
any device work."""
import numpy as np
import pytest

import oracle
from cases import build
from orctm_amd import abi, synth
from orctm_amd.nh import NHError, NHSolver, options

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("first, second", [("regional", "closed"), ("transect", "one_layer")])
def test_noalloc_apply_matches_apply_on_two_calls(first, second):
    print(f"[noalloc] {first} then {second}", flush=True)
    for k, name in enumerate((first, second)):
        g, divg = build(name)
        A, S = oracle.assemble(g)
        x = synth.normal(70 + k, 5, g.cells)
        with NHSolver(g, options(precision=abi.FP64)) as s:
            y_na = s.apply_noalloc(x)
            y = s.apply(x)
            y_na2 = s.apply_noalloc(2.0 * x)
        assert bits_equal(y_na, y), name
        assert bits_equal(y_na, oracle.matvec(g, A, x)), name
        assert bits_equal(y_na2, oracle.matvec(g, A, 2.0 * x)), name


def test_noalloc_apply_mixed_precision_handle():
    g, divg = build("cyclic_x")
    A, S = oracle.assemble(g)
    x = synth.normal(72, 5, g.cells)
    with NHSolver(g, options(precision=abi.MIXED)) as s:
        assert bits_equal(s.apply_noalloc(x), s.apply(x))


def test_noalloc_apply_refuses_systems_beyond_its_staging():
    # 256 x 256 x 130 = 8.5 M padded cells > 8388608: refused before any device work
    g = synth.build(synth.GridSpec(256, 256, 130, bathy="flat", depth=4000.0, seed=5), with_rhs=False)
    with NHSolver(g, options(precision=abi.FP64)) as s:
        with pytest.raises(NHError) as ei:
            s.apply_noalloc(np.zeros(g.cells))
        assert ei.value.code == abi.ERR_INVALID
        assert "staging limit" in abi.last_error()
