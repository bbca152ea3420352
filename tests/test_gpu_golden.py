"""GPU library (C ABI) against the golden fixtures: operator bit-exact with the independent dense
transliteration, solutions within 1e-7 of the FP64 direct solve, rel.res <= 1e-10."""
import numpy as np
import pytest

import golden_io
from orctm_amd import abi
from orctm_amd.nh import NHSolver, options

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.NAMES)
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_gpu_matches_golden(name, precision):
    g, divg, M, b, p = golden_io.load(name)
    print(f"[golden] {name} precision={precision}", flush=True)
    pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
    with NHSolver(g, options(precision=precision)) as s:
        A = s.operator()
        st = s.solve(divg, pnh)
    import oracle

    assert (oracle.to_scipy(g, A) != M).nnz == 0
    x = g.interior(pnh)
    assert st["status"] == "OK"
    assert np.linalg.norm(b - M @ x) / np.linalg.norm(b) <= 1e-10
    assert np.linalg.norm(x - p) / np.linalg.norm(p) <= 1e-7
