"""The oracle's interior-point iteration (oracle/ipm_ref.py) pinned by mathematics: it converges to the
known optimum of a 2 x 2 SDP and to the optimum of a small LP found by an independent LP solver
(scipy's HiGHS), its first iteration reproduces the closed forms of the residues at the SDPB starting
point, and the golden records (tests/golden/ipm_small.json) are what it computes."""
import mpmath as mp
import numpy as np
import pytest

from oracle import ipm_ref as IR
from tests import ipm_cases as C


def blocks_of(bl):
    return [IR.Block(b["dim"], b["K"], b["Ve"], b["Vo"], b["c"], b["B"], prec=1600) for b in bl]


def test_known_answer_sdp():
    bl, b, opt = C.two_by_two()
    log, res = IR.solve(blocks_of(bl), b, max_iter=300, prec=300, **C.QUICK)
    assert res["terminateReason"] == "found primal-dual optimal solution"
    assert abs(res["primalObjective"] - opt) < mp.mpf(10) ** -28          # gap < 1e-30 at termination
    assert abs(res["dualObjective"] - opt) < mp.mpf(10) ** -28


def test_lp_matches_linprog():
    from scipy.optimize import linprog
    bl, b, (c, Bm, bf) = C.small_lp()
    log, res = IR.solve(blocks_of(bl), b, max_iter=300, prec=300, **C.QUICK)
    assert res["terminateReason"] == "found primal-dual optimal solution"
    lp = linprog(c, A_eq=Bm.T, b_eq=bf, bounds=[(0, None)] * len(c), method="highs")
    assert lp.status == 0
    assert abs(float(res["primalObjective"]) - lp.fun) <= 1e-9 * max(1.0, abs(lp.fun))


def test_first_iteration_closed_forms():
    """at x = 0, y = 0, X = Y = 10^20 I: mu = 10^40, both objectives 0, P-err = 10^20, p-err = max |b|,
    D-err = max_p |c_p - 10^20 Tr(A_p)| with Tr(A_p) = [r == c] (|ve_k|^2 + |vo_k|^2), R-err = 0"""
    g = C.golden()
    bl, b = C.golden_blocks(g)
    log, _ = IR.solve(blocks_of(bl), b, max_iter=1, prec=1600)
    r = log[0]
    mp.mp.prec = 1600
    s = mp.mpf(10) ** 20
    assert r["mu"] == s * s and r["pobj"] == 0 and r["dobj"] == 0 and r["R_err"] == 0
    assert r["primal_err_P"] == s
    assert r["primal_err_p"] == max(abs(v) for v in b)
    worst = mp.mpf(0)
    for blk in bl:
        K = blk["K"]
        for s_, (cc, rr) in enumerate(IR.segments(blk["dim"])):
            for k in range(K):
                tr = mp.fsum(V[m][k] ** 2 for V in (blk["Ve"], blk["Vo"]) for m in range(len(V))) if cc == rr else 0
                worst = max(worst, abs(blk["c"][s_ * K + k] - s * tr))
    assert abs(r["dual_err"] - worst) <= mp.mpf(2) ** -1000 * worst


def test_golden_rows_reproduce():
    g = C.golden()
    bl, b = C.golden_blocks(g)
    log, _ = IR.solve(blocks_of(bl), b, max_iter=3, prec=1600)
    for r, gr in zip(log, g["rows"]):
        for k in C.KEYS:
            a, e = r[k], mp.mpf(gr[k])
            assert abs(a - e) <= mp.mpf(10) ** -38 * max(abs(a), abs(e), mp.mpf(10) ** -300), (r["iteration"], k)
