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


def test_tile_residues_kept_through_the_factorisation(monkeypatch):
    """the factorisation reuses the residues of its tiles' X and factor in the panel solves: the
    factor and tiles are those of a run that splits every tile operand again, with fewer splits,
    and no tile keeps residues afterwards"""
    from sdpvibe_amd import batched as Bm
    L, nb, n = 8, 2, 200
    m, e, _ = spd_blocks(nb, n, L, seed=17, uspread=12)
    A = F.MPF(cu(m), cu(e))
    calls = [0]
    desc = Bm._residue_desc

    def counted(*a):
        f = desc(*a)

        def g(*b):
            calls[0] += 1
            return f(*b)
        return g
    monkeypatch.setattr(Bm, "_residue_desc", counted)
    runs = []
    for budget in (Bm.TILE_RES_BUDGET, 0):
        monkeypatch.setattr(Bm, "TILE_RES_BUDGET", budget)
        calls[0] = 0
        C = F.Chol(A)
        runs.append((C, calls[0]))
        assert all(t.res is None for t in C.tiles)
    (C1, n1), (C0, n0) = runs
    assert n1 < n0, (n1, n0)
    assert torch.equal(C1.Lt.limbs, C0.Lt.limbs) and torch.equal(C1.k, C0.k)
    for t1, t0 in zip(C1.tiles, C0.tiles):
        assert torch.equal(t1.X.limbs, t0.X.limbs) and torch.equal(t1.Lmp.limbs, t0.Lmp.limbs)
