"""The GPU interior-point driver (sdpvibe_amd/ipm.py, the counterpart of sdpvibe/solver.py:680-982)
against the oracle's dense mpmath iteration (oracle/ipm_ref.py, pinned in test_ipm_oracle.py):

- the golden polynomial SDP: 20 iterations, every record's mu, objectives, errors, step lengths and
  centering parameter agree with the oracle's to 13 significant digits (errors below 1e-150 are
  rounding noise in both and only have to stay there), with the same verdict;
- the batched and the per-block layouts of an SDP give bitwise identical records and iterates;
- known answers solved to optimality: a 2 x 2 SDP (optimum -2) and a small LP (scipy's HiGHS).
"""
import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.ipm import Solver
from sdpvibe_amd.sdp import SDP
from tests import ipm_cases as C

pytestmark = pytest.mark.gpu
FIELDS = {"mu": "mu", "pobj": "P-obj", "dobj": "D-obj", "gap": "gap", "primal_err_P": "P-err",
          "primal_err_p": "p-err", "dual_err": "D-err", "R_err": "R-err", "alphaP": "P-step", "alphaD": "D-step",
          "beta": "beta"}
NOISE = {"primal_err_P", "primal_err_p", "dual_err", "R_err"}


def gpu_sdp(blocks, b, L, layout="batched"):
    """MPF SDP on the GPU from mpmath-valued blocks (values rounded to NB bits)"""
    out = []
    for bl in blocks:
        K = bl["K"]
        V = [None if not v else F.from_mp([x for r in v for x in r], L, (len(v), K)) for v in (bl["Ve"], bl["Vo"])]
        P, N = len(bl["c"]), len(bl["B"][0])
        out.append(dict(dim=bl["dim"], K=K, Ve=V[0], Vo=V[1], c=F.from_mp(bl["c"], L, (P, 1)),
                        B=F.from_mp([x for r in bl["B"] for x in r], L, (P, N))))
    return SDP.from_blocks(out, F.from_mp(b, L, (len(b), 1)), L, layout=layout)


def test_ipm_matches_oracle_golden():
    g = C.golden()
    L = g["L"]
    with mp.workprec(1700):
        blocks, b = C.golden_blocks(g)
    S = Solver(gpu_sdp(blocks, b, L), L)
    res = S.run(max_iter=len(g["rows"]))
    assert res["terminateReason"] == g["terminate"]
    assert len(S.log) == len(g["rows"])
    mp.mp.prec = 32 * L + 64
    worst = {}
    for row, ref in zip(S.log, g["rows"]):
        for k, name in FIELDS.items():
            a, e = mp.mpf(row[name]), mp.mpf(ref[k])
            if k in NOISE and abs(e) < mp.mpf(10) ** -150:
                assert abs(a) < mp.mpf(10) ** -150, (row["iteration"], k, a, e)
                continue
            d = abs(a - e) / max(abs(a), abs(e), mp.mpf(10) ** -300)
            worst[k] = max(worst.get(k, mp.mpf(0)), d)
            assert d <= mp.mpf(10) ** -13, (row["iteration"], k, mp.nstr(a, 20), mp.nstr(e, 20))
    print("\nworst relative difference vs the oracle over 20 iterations: "
          + ", ".join(f"{k} {mp.nstr(v, 2)}" for k, v in worst.items()))


def test_batched_equals_plain():
    sd = SDP.polynomial(n1=3, n3=2, degree=6, N=7, L=24, seed=3)
    sp = SDP.polynomial(n1=3, n3=2, degree=6, N=7, L=24, seed=3, layout="plain")
    assert len(sd.groups) == 2 and len(sp.groups) == 5
    A, B = Solver(sd, 24), Solver(sp, 24)
    A.run(max_iter=6)
    B.run(max_iter=6)
    for ra, rb in zip(A.log, B.log):
        for k in FIELDS.values():
            assert ra[k] == rb[k], (ra["iteration"], k)
    assert torch.equal(A.y.m, B.y.m) and torch.equal(A.y.e, B.y.e)


def test_graph_replays_equal_eager_iterations():
    """chol(Q) and the solves with its factor replayed from HIP graphs (iterations 2..6: five replays,
    eager work and host synchronisations in between) give bitwise the eager iterations"""
    sd = SDP.polynomial(n1=4, n3=2, degree=8, N=40, L=24, seed=5)
    A, B = Solver(sd, 24), Solver(sd, 24)
    A.graphs, B.graphs = True, False
    A.run(max_iter=6)
    B.run(max_iter=6)
    assert A._qgraph is not None and B._qgraph is None
    for ra, rb in zip(A.log, B.log):
        for k in FIELDS.values():
            assert ra[k] == rb[k], (ra["iteration"], k)
    assert torch.equal(A.y.m, B.y.m) and torch.equal(A.y.e, B.y.e)
    for sa, sb in zip(A.state, B.state):
        assert torch.equal(sa.X.m, sb.X.m) and torch.equal(sa.x.m, sb.x.m)


def test_known_answer_sdp():
    bl, b, opt = C.two_by_two()
    L = 8
    S = Solver(gpu_sdp(bl, b, L), L, **C.QUICK)
    res = S.run(max_iter=300)
    assert res["terminateReason"] == "found primal-dual optimal solution"
    assert abs(res["primalObjective"] - opt) < mp.mpf(10) ** -28          # gap < 1e-30 at termination
    assert abs(res["dualObjective"] - opt) < mp.mpf(10) ** -28


def test_lp_matches_linprog():
    from scipy.optimize import linprog
    bl, b, (c, Bm, bf) = C.small_lp()
    L = 8
    S = Solver(gpu_sdp(bl, b, L), L, **C.QUICK)
    res = S.run(max_iter=300)
    assert res["terminateReason"] == "found primal-dual optimal solution"
    lp = linprog(c, A_eq=Bm.T, b_eq=bf, bounds=[(0, None)] * len(c), method="highs")
    assert abs(float(res["primalObjective"]) - lp.fun) <= 1e-9 * max(1.0, abs(lp.fun))


def test_polynomial_sdp_is_strictly_feasible():
    """X* = sum_p A_p x*_p = I for x* = 1 on the diagonal segments (orthonormal bases), and B^T x* = b"""
    L = 24
    sd = SDP.polynomial(n1=2, n3=1, degree=8, N=6, L=L, seed=5)
    S = Solver(sd, L)
    mp.mp.prec = 32 * L + 64
    for g in sd.groups:
        xs = F.zeros((g.nb, g.P, 1), L)
        one = F.scalar(1, L)
        from sdpvibe_amd.sdp import segments
        for c, r, s in segments(g.dim):
            if c == r:
                xs.m[:, :, s * g.K:(s + 1) * g.K] = one.m[:, :, None, None]
                xs.e[:, s * g.K:(s + 1) * g.K] = one.e[0]
        Xs = S.constraint_sum(g, xs)
        I = F.eye(g.n, L, 1)
        for z in range(g.nbp):
            D = F.add(F.MPF(Xs.m[:, z].contiguous(), Xs.e[z].contiguous()), I, -1)
            D = F.MPF(torch.where(g.mask2[z][None], D.m, 0), torch.where(g.mask2[z], D.e, F.EZ))
            assert F.max_abs(D) < mp.mpf(2) ** -(32 * L - 40)
