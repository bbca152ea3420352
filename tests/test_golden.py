"""Golden fixtures (tests/golden/*.npz): inputs in Fortran layout, the operator from an independent
dense transliteration of compute_matrix, and FP64 direct solutions. CPU checks of the oracle here;
the GPU library is checked against the same fixtures in test_gpu_golden.py."""
import numpy as np
import pytest

import golden_io
import oracle


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_oracle_operator_equals_golden(name):
    g, divg, M, b, p = golden_io.load(name)
    A, S = oracle.assemble(g)
    Mo = oracle.to_scipy(g, A)
    assert (Mo != M).nnz == 0


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_oracle_rhs_and_solve_match_golden(name):
    g, divg, M, b, p = golden_io.load(name)
    assert np.array_equal(oracle.rhs(g, divg), b)
    A, S = oracle.assemble(g)
    for prec in (0, 1):
        x, st = oracle.solve(g, A, S, b, precision=prec)
        assert st.status == 0
        assert np.linalg.norm(x - p) / np.linalg.norm(p) <= 1e-7
    x, st = oracle.fgmres(g, A, S, b)
    assert np.linalg.norm(x - p) / np.linalg.norm(p) <= 1e-7


def test_fixtures_present():
    assert len(golden_io.NAMES) >= 3
