"""Objective sensitivities at the optimum (sdpvibe_amd/adjoint.py) on a small polynomial SDP at 1086 bits
(L = 34), solved to a duality gap below 1e-70:

- the gradients of the primal and the dual objective through the adjoint Schur-complement solve (the
  torch.autograd path) equal the first-order perturbation estimate dv = dc.x + db.y - x^T dB y for 64
  random unit directions, and both agree with central finite differences (step 1e-40, each perturbed
  SDP solved to the same gap) to 20 significant digits;
- the adjoint solve is the transpose of the forward (tangent) solve of the same Newton system, to
  working precision.
"""
import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.adjoint import (SDPValue, adjoint, contract, directional, objective_gradients, perturbed,
                                  primal_vector, refactor, tangent)
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
    """64 directions: the gradients of the primal and of the dual objective through the adjoint solve
    (the torch.autograd path) equal the first-order estimate; two of them against central differences"""
    sdp = SDP.polynomial(n1=2, n3=1, degree=4, N=5, L=L, seed=1)
    base = solve(sdp)
    mp.mp.prec = 32 * L + 64
    K = 64
    dc, dB, db = directions(sdp, K, seed=2)
    est = directional(base, dc, dB, db)
    refactor(base)
    for which in ("primal", "dual"):
        adj = contract(objective_gradients(base, which), dc, dB, db)
        worst = max(abs(a - e) / abs(e) for a, e in zip(adj, est))
        print(f"\n{which} objective: adjoint vs first-order estimate over {K} directions, worst relative "
              f"difference {mp.nstr(worst, 3)}")
        assert worst < mp.mpf(10) ** -40
    # the same numbers through torch.autograd
    theta = torch.zeros(K, dtype=torch.float64, requires_grad=True)
    SDPValue.apply(theta, base, (dc, dB, db), "primal", "adjoint").backward()
    assert np.allclose(theta.grad.numpy(), [float(v) for v in est], rtol=1e-15, atol=0)
    h = mp.mpf(10) ** -40
    for k in range(2):
        vp = solve(perturbed(sdp, dc[k], dB[k], db[k], h)).result["primalObjective"]
        vm = solve(perturbed(sdp, dc[k], dB[k], db[k], -h)).result["primalObjective"]
        fd = (vp - vm) / (2 * h)
        rel = abs(fd - est[k]) / abs(est[k])
        print(f"direction {k}: first-order {mp.nstr(est[k], 25)}  central difference {mp.nstr(fd, 25)}  "
              f"relative difference {mp.nstr(rel, 3)}")
        assert rel < mp.mpf(10) ** -20


def _pair(A, B):
    """sum of the entrywise products of two MPF arrays of one shape, as an mpmath number"""
    p = F.mul(A, B)
    return F.total(F.MPF(p.m.reshape(p.L, -1), p.e.reshape(-1)))


@pytest.mark.timeout(900)
def test_adjoint_is_the_transposed_newton_solve():
    """adjoint(gx, gy) contracted with a direction equals gx . dx + gy . dy for the forward (tangent)
    solve of the same Newton system, to working precision; and the tangent of the solution follows
    central differences of re-solved SDPs to the accuracy the final iterate's linearisation has"""
    sdp = SDP.polynomial(n1=2, n3=1, degree=4, N=5, L=L, seed=1)
    base = solve(sdp)
    mp.mp.prec = 32 * L + 64
    dc, dB, db = directions(sdp, 2, seed=3)
    x, y = primal_vector(base), base.y
    rng = np.random.default_rng(4)
    gx = _mpf_from_float(torch.from_numpy(rng.standard_normal((x.shape[0], 1))).cuda(), L)
    gy = _mpf_from_float(torch.from_numpy(rng.standard_normal((y.shape[0], 1))).cuda(), L)
    refactor(base)
    grads = adjoint(base, gx, gy)
    est = contract(grads, dc, dB, db)
    for k in range(2):
        dx, dy = tangent(base, dc[k], dB[k], db[k])
        fwd = _pair(gx, dx) + _pair(gy, dy)
        rel = abs(fwd - est[k]) / abs(est[k])
        print(f"\ndirection {k}: adjoint {mp.nstr(est[k], 25)}  tangent {mp.nstr(fwd, 25)}  rel {mp.nstr(rel, 3)}")
        assert rel < mp.mpf(10) ** -150
    # the Newton system of the final iterate (gap 1e-70) is the derivative of the central-path point
    # there, which approximates the derivative of the computed optimum to about 1e-16 on this SDP
    # (the optimal value, stationary, agrees to 1e-50: test above)
    h = mp.mpf(10) ** -25
    ls = []
    for t in (h, -h):
        s = solve(perturbed(sdp, dc[0], dB[0], db[0], t))
        ls.append(_pair(gx, primal_vector(s)) + _pair(gy, s.y))
    fd = (ls[0] - ls[1]) / (2 * h)
    rel = abs(fd - est[0]) / abs(est[0])
    print(f"functional of (x, y): adjoint {mp.nstr(est[0], 25)}  central difference {mp.nstr(fd, 25)}  rel {mp.nstr(rel, 3)}")
    assert rel < mp.mpf(10) ** -12
