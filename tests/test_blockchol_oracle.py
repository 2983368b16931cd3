"""The oracle's tiled Cholesky (sdpvibe/blockchol.py) pinned by mathematics: solve(B) = Q^-1 B to
working precision (mpmath), and the tiled factor agrees with the monolithic one (oracle Chol) to
working precision, as the reference states for its own (different rounding order)."""
import random

import mpmath as mpm
import numpy as np
import pytest

from oracle import blockchol_ref as BC
from tests.intutil import from_planes, mpf_host, spd_blocks


def _mat(m, e):
    L, R, C = m.shape
    ints = np.array(from_planes(m.reshape(L, -1)), dtype=object).reshape(R, C)
    return mpm.matrix([[mpm.ldexp(mpm.mpf(int(ints[i, j])), int(e[i, j])) if ints[i, j] else mpm.mpf(0)
                        for j in range(C)] for i in range(R)])


@pytest.mark.parametrize("L,N,b", [(3, 70, 32), (4, 64, 40)])
def test_blockchol_solve(L, N, b):
    m, e, _ = spd_blocks(1, N, L, seed=N + b, uspread=6)
    Q = (m[:, 0], e[0])
    bc = BC.BlockChol(Q, b)
    rng = random.Random(N)
    bm, be = mpf_host([(rng.randrange(-(1 << 80), 1 << 80), rng.randrange(-20, 20)) for _ in range(N * 2)], L)
    B = (bm.reshape(L, N, 2), be.reshape(N, 2))
    X = bc.solve_upper(bc.solve_lower(B))
    mpm.mp.prec = 64 * L + 200
    Qm, Bm, Xm = _mat(*Q), _mat(*B), _mat(*X)
    cond = mpm.mnorm(Qm, 1) * mpm.mnorm(mpm.inverse(Qm), 1)
    NB = 32 * L - 2
    for c in range(2):
        xe = mpm.lu_solve(Qm, mpm.matrix([Bm[i, c] for i in range(N)]))
        xn = max(abs(xe[i]) for i in range(N))
        err = max(abs(Xm[i, c] - xe[i]) for i in range(N))
        assert err <= cond * mpm.ldexp(1, 48 - NB) * N * N * xn, float(mpm.log(err / xn, 2))
