"""CPU check of the restated CPU-baseline generator (tests/test_gpu_golden.py): in float64, the
Cholesky of A = G G^T + n I and L^-1 B reproduce the reference's printed check values
(tests/golden/mpf_baseline_check.json) to 7 significant digits, so the GPU test compares against
the right matrices."""
import numpy as np
import pytest

from tests.test_gpu_golden import CASES, _lcg_values


@pytest.mark.parametrize("case", [c for c in CASES if c["n"] <= 1024], ids=lambda c: c["file"])
def test_generator_reproduces_check_values(case):
    n, R = case["n"], case["nrhs"]
    g, s = _lcg_values(n * n)
    b, _ = _lcg_values(n * R, s)
    G = np.array([a / d for a, d in g]).reshape(n, n)
    bcol = np.array([b[i * R][0] / b[i * R][1] for i in range(n)])
    A = G @ G.T + n * np.eye(n)
    Lf = np.linalg.cholesky(A)
    p = np.linalg.solve(Lf, bcol)
    assert abs(Lf[n - 1, n - 1] / float(case["L_last"]) - 1) < 6e-7
    assert abs(p[n - 1] / float(case["P_last_0"]) - 1) < 6e-7
