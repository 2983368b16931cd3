"""The oracle's single-matrix Cholesky / triangular-solve restatement (sdpvibe/linalg.py,
sdpvibe/basecase.py:259-332) pinned by mathematics (mpmath at high precision): backward error of
the factor, X L = I for the inverse, and the residual of forward substitution and triangular
solves, each within a bound in bits that a wrong index, sign or transposition would break.
The GPU path is compared with this restatement bit for bit in test_gpu_linalg.py."""
import random

import mpmath as mpm
import numpy as np
import pytest

from oracle import linalg_ref as LR
from tests.intutil import from_planes, spd_tile, to_planes


def _vals(M):
    """OM -> mpmath matrix of its values"""
    L, R, C = M.limbs.shape
    ints = np.array(from_planes(M.limbs.reshape(L, -1)), dtype=object).reshape(R, C)
    ev = M.expv()
    return mpm.matrix([[mpm.ldexp(mpm.mpf(int(ints[i, j])), int(ev[j])) for j in range(C)] for i in range(R)])


def _spd(n, L, logc, seed):
    F = 32 * L - 40
    A = spd_tile(n, logc, seed, F)
    return LR.OM(to_planes([v for row in A for v in row], L).reshape(L, n, n), -F)


def _maxabs(M):
    return max(abs(M[i, j]) for i in range(M.rows) for j in range(M.cols))


@pytest.mark.parametrize("L,n,logc", [(3, 40, 4), (4, 64, 8), (3, 96, 6)])
def test_chol_inv_oracle(L, n, logc):
    A = _spd(n, L, logc, seed=n + L)
    Lf, X = LR.chol_inv(A)
    mpm.mp.prec = 64 * L + 200
    F = 32 * L - 40
    Am, Lm, Xm = _vals(A), _vals(Lf), _vals(X)
    assert all(Lm[i, j] == 0 for i in range(n) for j in range(i + 1, n))
    bw = _maxabs(Lm * Lm.T - Am)
    assert bw == 0 or mpm.log(bw, 2) <= -F + 10, float(mpm.log(bw, 2))
    xl = _maxabs(Xm * Lm - mpm.eye(n))
    cond_bits = float(mpm.log(_maxabs(Xm), 2))
    assert xl == 0 or mpm.log(xl, 2) <= -F + 2 * cond_bits + 12, (float(mpm.log(xl, 2)), cond_bits)


@pytest.mark.parametrize("L,leaf,refine", [(3, "subst", 0), (3, "gemm", 1), (4, "subst", 0)])
def test_chol_trsm_oracle(L, leaf, refine):
    n, m = 64, 5
    A = _spd(n, L, 6, seed=L)
    Lf, tiles = LR.chol(A, tile_inverses=(leaf == "gemm"), panel_leaf=leaf, refine=refine)
    mpm.mp.prec = 64 * L + 200
    F = 32 * L - 40
    Am, Lm = _vals(A), _vals(Lf)
    bw = _maxabs(Lm * Lm.T - Am)
    assert bw == 0 or mpm.log(bw, 2) <= -F + 10, float(mpm.log(bw, 2))
    rng = random.Random(L)
    Bl = to_planes([rng.randrange(-(1 << (32 * L - 4)), 1 << (32 * L - 4)) for _ in range(n * m)], L).reshape(L, n, m)
    be = np.array([rng.randrange(-3000, -2900) for _ in range(m)], np.int64)
    B = LR.OM(Bl, be)
    P = LR.trsm_left(Lf, B, tiles, leaf=leaf, refine=refine)
    Pm, Bm = _vals(P), _vals(B)
    for c in range(m):
        r = max(abs(sum(Lm[i, k] * Pm[k, c] for k in range(i + 1)) - Bm[i, c]) for i in range(n))
        bn = max(abs(Bm[i, c]) for i in range(n))
        assert r <= bn * mpm.ldexp(1, -F + 16), float(mpm.log(r / bn, 2))


@pytest.mark.parametrize("L", [2, 4])
def test_tile_trsv_oracle(L):
    """forward substitution on the factored tile: L y = v to the wide precision"""
    b, nvec = 32, 3
    A = _spd(b, L, 5, seed=11 * L)
    _, t = LR.chol_tile(A)
    rng = random.Random(L)
    V = to_planes([rng.randrange(-(1 << (32 * L - 41)), 1 << (32 * L - 41)) for _ in range(nvec * b)], L)
    V = V.reshape(L, nvec, b)
    ve = np.array([-(32 * L - 40)] * nvec, np.int64)
    Y, ye = LR.trsv_tile(t, V, ve)
    mpm.mp.prec = 64 * L + 200
    Lm = _vals(t.Lmp)
    F = 32 * L - 40
    for q in range(nvec):
        vq = from_planes(V[:, q])
        yq = from_planes(Y[:, q])
        vv = [mpm.ldexp(v, int(ve[q])) for v in vq]
        yy = [mpm.ldexp(y, int(ye[q])) for y in yq]
        r = max(abs(sum(Lm[i, k] * yy[k] for k in range(i + 1)) - vv[i]) for i in range(b))
        yn = max(abs(y) for y in yy)
        assert r <= yn * mpm.ldexp(1, -F + 12), float(mpm.log(r / yn, 2))
