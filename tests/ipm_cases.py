"""SDPs with known answers and the golden polynomial SDP, shared by the oracle's CPU tests and the GPU
driver's tests. Plain Python/mpmath data: each block is dict(dim, K, Ve, Vo, c, B) of mpmath numbers
(Ve, Vo lists of rows or None)."""
import json
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# known-answer problems start near their solution scale: fewer iterations to optimality
QUICK = dict(initialMatrixScalePrimal="10", initialMatrixScaleDual="10")
KEYS = ["mu", "pobj", "dobj", "gap", "primal_err_P", "primal_err_p", "dual_err", "R_err", "alphaP", "alphaD", "beta"]


def two_by_two():
    """min x10 s.t. X = [[x00, x10/2], [x10/2, x11]] >= 0, x00 = 1, x11 = 1: optimum -2 (X = [[1, -1], [-1, 1]])"""
    one = [[mp.mpf(1)]]
    return [dict(dim=2, K=1, Ve=one, Vo=None, c=[0, 1, 0], B=[[1, 0], [0, 0], [0, 1]])], [1, 1], mp.mpf(-2)


def small_lp(n=8, m=3, seed=1):
    """min c.x s.t. x >= 0, B^T x = b as an SDP of 1 x 1 blocks (K = 1, one basis row); feasible and
    bounded by construction. Returns (blocks, b, (c, B, b) as float64 arrays for an LP solver)."""
    rng = np.random.default_rng(seed)
    Bm = rng.standard_normal((n, m))
    b = Bm.T @ (rng.random(n) + 0.1)
    c = rng.random(n) + Bm @ rng.standard_normal(m)
    one = [[mp.mpf(1)]]
    blocks = [dict(dim=1, K=1, Ve=one, Vo=None, c=[c[i]], B=[list(Bm[i])]) for i in range(n)]
    return blocks, list(b), (c, Bm, b)


def _val(me):
    """exact mpmath value of an MPF entry stored as (mantissa, exponent)"""
    m, e = int(me[0]), me[1]
    if m == 0:
        return mp.mpf(0)
    with mp.workprec(max(64, abs(m).bit_length() + 8)):
        return mp.ldexp(mp.mpf(m), e)


def golden():
    """the polynomial SDP of tests/golden/ipm_small.json (exact MPF values as (mantissa, exponent)) and
    the oracle's rows"""
    with open(os.path.join(HERE, "golden", "ipm_small.json")) as f:
        g = json.load(f)
    return g


def golden_blocks(g):
    """mpmath values of the golden SDP"""
    blocks = [dict(dim=bl["dim"], K=bl["K"], Ve=[[_val(v) for v in r] for r in bl["Ve"]],
                   Vo=[[_val(v) for v in r] for r in bl["Vo"]], c=[_val(v) for v in bl["c"]],
                   B=[[_val(v) for v in r] for r in bl["B"]]) for bl in g["blocks"]]
    return blocks, [_val(v) for v in g["b"]]
