"""Objective sensitivities at the optimum (sdpvibe_amd/adjoint.py) on a small polynomial SDP at 1086 bits
(L = 34), solved to a duality gap below 1e-70: the first-order perturbation estimate
dv = dc.x + db.y - x^T dB y, evaluated for a batch of random unit directions with exact GEMMs and
through torch.autograd, agrees with central finite differences (step 1e-40, each perturbed SDP solved to
the same gap) to 20 significant digits."""
import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.adjoint import SDPValue, adjoint, directional, perturbed, primal_vector, refactor
from sdpvibe_amd.ipm import Solver
from sdpvibe_amd.sdp import SDP
from sdpvibe_amd.steplen import _mpf_from_float

pytestmark = pytest.mark.gpu
L = 34
TIGHT = dict(dualityGapThreshold="1e-70", primalErrorThreshold="1e-70", dualErrorThreshold="1e-70",
             initialMatrixScalePrimal="1e4", initialMatrixScaleDual="1e4")


def solve(sdp):
    s = Solver(sdp, L, **TIGHT)
    res = s.run(max_iter=400)
    assert res["terminateReason"] == "found primal-dual optimal solution", res["terminateReason"]
    return s


def directions(sdp, K, seed):
    """K Gaussian directions of unit Euclidean norm over (c, B, b) jointly"""
    rng = np.random.default_rng(seed)
    P = sum(g.nb * g.P for g in sdp.groups)
    N = sdp.N
    raw = rng.standard_normal((K, P + P * N + N))
    raw /= np.linalg.norm(raw, axis=1, keepdims=True)
    t = lambda a: _mpf_from_float(torch.from_numpy(np.ascontiguousarray(a)).cuda(), L)
    return t(raw[:, :P].reshape(K, P, 1)), t(raw[:, P:P + P * N].reshape(K, P, N)), t(raw[:, P + P * N:].reshape(K, N, 1))


@pytest.mark.timeout(900)
def test_objective_sensitivity_vs_finite_differences():
    sdp = SDP.polynomial(n1=2, n3=1, degree=4, N=5, L=L, seed=1)
    base = solve(sdp)
    mp.mp.prec = 32 * L + 64
    K = 8
    dc, dB, db = directions(sdp, K, seed=2)
    est = directional(base, dc, dB, db)
    # the same numbers through torch.autograd
    theta = torch.zeros(K, dtype=torch.float64, requires_grad=True)
    SDPValue.apply(theta, base, (dc, dB, db)).backward()
    assert np.allclose(theta.grad.numpy(), [float(v) for v in est], rtol=1e-15, atol=0)
    h = mp.mpf(10) ** -40
    for k in range(2):
        vp = solve(perturbed(sdp, dc[k], dB[k], db[k], h)).result["primalObjective"]
        vm = solve(perturbed(sdp, dc[k], dB[k], db[k], -h)).result["primalObjective"]
        fd = (vp - vm) / (2 * h)
        rel = abs(fd - est[k]) / abs(est[k])
        print(f"\ndirection {k}: first-order {mp.nstr(est[k], 25)}  central difference {mp.nstr(fd, 25)}  "
              f"relative difference {mp.nstr(rel, 3)}")
        assert rel < mp.mpf(10) ** -20
