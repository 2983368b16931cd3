"""The oracle's batched Cholesky / triangular-solve composition (sdpvibe/batched.py:243-352,
sdpvibe/mpf.py:461-532) pinned by mathematics: backward error of the factor, residuals of the
solves and forward errors against mpmath at high precision, stated in bits. The GPU path is
compared with this composition bit for bit in test_gpu_chol.py."""
import mpmath as mpm
import numpy as np
import pytest

from oracle import batched_ref as BR
from tests.intutil import EZ, from_planes, mpf_host, spd_blocks


def _mpf_matrix(planes, exps, shape):
    ints = from_planes(planes.reshape(planes.shape[0], -1))
    ex = exps.reshape(-1)
    vals = [mpm.mpf(0) if v == 0 else mpm.ldexp(mpm.mpf(v), int(x)) for v, x in zip(ints, ex)]
    return np.array(vals, dtype=object).reshape(shape)


def _exact(exact_b):
    n = len(exact_b)
    return mpm.matrix([[mpm.ldexp(mpm.mpf(v), ev) for (v, ev) in exact_b[i]] for i in range(n)])


@pytest.mark.parametrize("L,nb,n,u", [(3, 2, 40, 8), (4, 1, 64, 20), (2, 3, 17, 3), (3, 1, 96, 12)])
def test_chol_factor_backward_error(L, nb, n, u):
    m, e, exact = spd_blocks(nb, n, L, seed=L * 100 + n, uspread=u)
    C = BR.Chol(m, e)
    mpm.mp.prec = 64 * L + 200
    F = 32 * L - 40
    Lt = from_planes(C.Lt.limbs.reshape(L, -1))
    npad = C.npad
    Lt = np.array(Lt, dtype=object).reshape(nb, npad, npad)
    for b in range(nb):
        assert all(Lt[b, i, j] == 0 for i in range(npad) for j in range(i + 1, npad))
        assert all(Lt[b, i, i] > 0 for i in range(npad))
        k = C.k[b]
        # equilibrated matrix from the rounded MPF input, exactly
        Am = _mpf_matrix(m[:, b], e[b], (n, n))
        worst = mpm.mpf(0)
        for i in range(n):
            for j in range(i + 1):
                dad = Am[i, j] * mpm.ldexp(1, -int(k[i]) - int(k[j]))
                llt = mpm.fsum(mpm.mpf(int(Lt[b, i, t])) * int(Lt[b, j, t]) for t in range(j + 1))
                worst = max(worst, abs(llt * mpm.ldexp(1, -2 * F) - dad))
        # |Lt Lt^T - D A D| <= n 2^(8-F): a handful of roundings per entry at 2^-F
        assert worst == 0 or mpm.log(worst, 2) <= -F + 8 + mpm.log(n, 2), float(mpm.log(worst, 2))
        # the identity padding stays exactly identity
        for i in range(n, npad):
            assert Lt[b, i, i] == 1 << F


@pytest.mark.parametrize("L,nb,n,r", [(3, 2, 40, 3), (4, 1, 64, 2)])
def test_chol_solve_residual(L, nb, n, r):
    m, e, exact = spd_blocks(nb, n, L, seed=7 * L + n, uspread=10)
    C = BR.Chol(m, e)
    import random
    rng = random.Random(n)
    vals = [(rng.randrange(-(1 << 90), 1 << 90), rng.randrange(-40, 40)) for _ in range(nb * n * r)]
    bm, be = mpf_host(vals, L)
    bm = bm.reshape(L, nb, n, r)
    be = be.reshape(nb, n, r)
    ym, ye = C.solve_lower(bm, be)
    xm, xe = C.solve_upper(ym, ye)
    mpm.mp.prec = 64 * L + 200
    NB = 32 * L - 2
    for b in range(nb):
        A = _mpf_matrix(m[:, b], e[b], (n, n))
        Bv = _mpf_matrix(bm[:, b], be[b], (n, r))
        X = _mpf_matrix(xm[:, b], xe[b], (n, r))
        Ama = mpm.matrix(A.tolist())
        cond = mpm.mnorm(Ama, 1) * mpm.mnorm(mpm.inverse(Ama), 1)
        for c in range(r):
            Xe = mpm.lu_solve(Ama, mpm.matrix([Bv[i, c] for i in range(n)]))
            xn = max(abs(Xe[i]) for i in range(n))
            err = max(abs(X[i, c] - Xe[i]) for i in range(n))
            # forward error <= cond * 2^(40 - NB) * n^2 (fixed point at 32L-40 bits, normwise)
            bound = cond * mpm.ldexp(1, 48 - NB) * n * n * xn
            assert err <= bound, (float(mpm.log(err / xn, 2)), float(mpm.log(cond, 2)))


def test_chol_diagonal_known_answer():
    """diagonal A = diag(4^j * c_j): k_j = floor((top + 1) / 2), Lt_jj = sqrt(c_j 4^j 2^-2k)"""
    L, n = 3, 5
    cs = [1, 2, 3, 4, 9]
    vals = []
    for i in range(n):
        for j in range(n):
            vals.append((cs[i] << (2 * i), 0) if i == j else (0, 0))
    m, e = mpf_host(vals, L)
    m = m.reshape(L, 1, n, n)
    e = e.reshape(1, n, n)
    C = BR.Chol(m, e)
    F = 32 * L - 40
    Lt = from_planes(C.Lt.limbs.reshape(L, -1))
    Lt = np.array(Lt, dtype=object).reshape(32, 32)
    for i in range(n):
        a = cs[i] << (2 * i)
        k = (a.bit_length() + 1) // 2          # top = bit length (value < 2^top)
        assert C.k[0, i] == k
        exact = mpm.sqrt(mpm.mpf(a) * mpm.ldexp(1, -2 * k))
        assert abs(mpm.mpf(int(Lt[i, i])) * mpm.ldexp(1, -F) - exact) <= mpm.ldexp(1, 4 - F)
        assert all(Lt[i, j] == 0 for j in range(n) if j != i)


def test_chol_not_pd_raises():
    L, n = 2, 3
    vals = [((1 if i == j else 0) * (-1 if i == 1 else 1), 0) for i in range(n) for j in range(n)]
    m, e = mpf_host(vals, L)
    with pytest.raises(ValueError):
        BR.Chol(m.reshape(L, 1, n, n), e.reshape(1, n, n))
    vals = [((1 if (i == j and i != 2) else 0), 0) for i in range(n) for j in range(n)]
    m, e = mpf_host(vals, L)
    assert (e.reshape(n, n)[2, 2] == EZ)
    with pytest.raises(ValueError):
        BR.Chol(m.reshape(L, 1, n, n), e.reshape(1, n, n))
