"""GPU batched Cholesky and solves (sdpvibe/mpf.py:461-532, sdpvibe/batched.py:243-352) bit for bit
against the oracle's composition of the same recursion (pinned in test_chol_oracle.py)."""
import random

import numpy as np
import pytest
import torch

from oracle import batched_ref as BR
from sdpvibe_amd import mpf as F
from tests.intutil import mpf_host, spd_blocks

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def chol_parity(L, nb, n, r, u, seed, two_d=False):
    m, e, _ = spd_blocks(nb, n, L, seed=seed, uspread=u)
    ref = BR.Chol(m, e)
    A = F.MPF(cu(m[:, 0]), cu(e[0])) if two_d else F.MPF(cu(m), cu(e))
    C = F.Chol(A)
    _eq(C.k.cpu().numpy().reshape(nb, n), ref.k, "k")
    _eq(C.Lt.limbs.cpu().numpy(), ref.Lt.limbs, "Lt")
    assert len(C.tiles) == len(ref.tiles)
    for i, (t, rt) in enumerate(zip(C.tiles, ref.tiles)):
        _eq(t.X.limbs.cpu().numpy(), rt.X.limbs, f"tile {i} X")
        _eq(t.X.exp.cpu().numpy(), rt.X.exp, f"tile {i} X exp")
        _eq(t.Lmp.limbs.cpu().numpy(), rt.Lmp.limbs, f"tile {i} L")
    rng = random.Random(seed)
    vals = [(rng.randrange(-(1 << 200), 1 << 200), rng.randrange(-60, 60)) for _ in range(nb * n * r)]
    vals[0] = (0, 0)
    bm, be = mpf_host(vals, L)
    bm, be = bm.reshape(L, nb, n, r), be.reshape(nb, n, r)
    B = F.MPF(cu(bm[:, 0]), cu(be[0])) if two_d else F.MPF(cu(bm), cu(be))
    Y = C.solve_lower(B)
    ym, ye = ref.solve_lower(bm, be)
    X = C.solve_upper(Y)
    xm, xe = ref.solve_upper(ym, ye)
    sq = (lambda a: a[:, 0]) if two_d else (lambda a: a)
    sqe = (lambda a: a[0]) if two_d else (lambda a: a)
    _eq(Y.m.cpu().numpy(), sq(ym), "solve_lower m")
    _eq(Y.e.cpu().numpy(), sqe(ye), "solve_lower e")
    _eq(X.m.cpu().numpy(), sq(xm), "solve_upper m")
    _eq(X.e.cpu().numpy(), sqe(xe), "solve_upper e")


@pytest.mark.parametrize("L,nb,n,r,u", [(2, 2, 40, 3, 5), (4, 3, 17, 1, 20), (8, 2, 64, 4, 30), (24, 2, 96, 3, 20),
                                        (24, 1, 130, 2, 40)])
def test_chol_parity(L, nb, n, r, u):
    chol_parity(L, nb, n, r, u, seed=L * 1000 + n)


def test_chol_parity_2d():
    chol_parity(24, 1, 50, 2, 10, seed=3, two_d=True)


def test_chol_not_pd():
    L, n = 3, 4
    vals = [((1 if i == j else 0) * (-1 if i == 2 else 1), 0) for i in range(n) for j in range(n)]
    m, e = mpf_host(vals, L)
    with pytest.raises(ValueError):
        F.Chol(F.MPF(cu(m.reshape(L, n, n)), cu(e.reshape(n, n))))
