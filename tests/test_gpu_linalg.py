"""GPU single-matrix Cholesky / triangular solves and the tile API (sdpvibe/linalg.py,
sdpvibe/basecase.py:259-332) bit for bit against the oracle's restatement (pinned in
test_linalg_oracle.py)."""
import random

import numpy as np
import pytest
import torch

from oracle import linalg_ref as LR
from sdpvibe_amd import basecase as BC
from sdpvibe_amd import linalg as LA
from sdpvibe_amd.mpmat import MP
from tests.intutil import spd_tile, to_planes

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def _same(M, R, what):
    _eq(M.limbs.cpu().numpy(), R.limbs, what + " limbs")
    ge = M.exp.cpu().numpy() if torch.is_tensor(M.exp) else np.full(M.shape[1], M.exp)
    _eq(ge, R.expv(), what + " exps")


def _spd(n, L, logc, seed):
    F = 32 * L - 40
    A = spd_tile(n, logc, seed, F)
    lim = to_planes([v for row in A for v in row], L).reshape(L, n, n)
    return MP(cu(lim), -F), LR.OM(lim, -F)


def _rhs(L, n, m, seed):
    rng = random.Random(seed)
    Bl = to_planes([rng.randrange(-(1 << (32 * L - 4)), 1 << (32 * L - 4)) for _ in range(n * m)], L).reshape(L, n, m)
    be = np.array([rng.randrange(-3000, -2900) for _ in range(m)], np.int64)
    return MP(cu(Bl), cu(be)), LR.OM(Bl, be)


def linalg_chol_inv_case(L, n, logc):
    A, Ar = _spd(n, L, logc, seed=n + L)
    Lf, X = LA.chol_inv(A)
    Lr, Xr = LR.chol_inv(Ar)
    _same(Lf, Lr, "chol_inv L")
    _same(X, Xr, "chol_inv X")
    B, Br = _rhs(L, n, 3, seed=L)
    _same(LA.trsm(X, B), LR.trsm(Xr, Br), "trsm")


def linalg_chol_case(L, n, leaf, refine, m=4):
    A, Ar = _spd(n, L, 6, seed=L + n)
    Lf, tiles = LA.chol(A, tile_inverses=(leaf == "gemm"), panel_leaf=leaf, refine=refine)
    Lr, tr = LR.chol(Ar, tile_inverses=(leaf == "gemm"), panel_leaf=leaf, refine=refine)
    _same(Lf, Lr, "chol L")
    B, Br = _rhs(L, n, m, seed=L + 1)
    _same(LA.trsm_left(Lf, B, tiles, leaf=leaf, refine=refine),
          LR.trsm_left(Lr, Br, tr, leaf=leaf, refine=refine), "trsm_left")


@pytest.mark.parametrize("L,n,logc", [(3, 40, 4), (4, 64, 8), (24, 96, 20), (24, 150, 40)])
def test_chol_inv(L, n, logc):
    linalg_chol_inv_case(L, n, logc)


@pytest.mark.parametrize("L,n,leaf,refine", [(3, 64, "subst", 0), (3, 64, "gemm", 1), (24, 96, "subst", 0),
                                             (24, 128, "gemm", 0), (8, 32, "subst", 0)])
def test_chol_trsm(L, n, leaf, refine):
    linalg_chol_case(L, n, leaf, refine)


@pytest.mark.parametrize("L,nvec", [(2, 1), (4, 9), (24, 40)])
def test_trsv_tile(L, nvec):
    b = 32
    A, Ar = _spd(b, L, 5, seed=11 * L)
    _, t = BC.chol_tile(A)
    _, tr = LR.chol_tile(Ar)
    _eq(t.Lw.cpu().numpy().view(np.uint32), tr.Lw, "tile Lw")
    _eq(t.RINV.cpu().numpy().view(np.uint32), tr.RINV, "tile RINV")
    rng = random.Random(L)
    V = to_planes([rng.randrange(-(1 << (32 * L - 41)), 1 << (32 * L - 41)) for _ in range(nvec * b)], L)
    V = V.reshape(L, nvec, b)
    ve = np.array([rng.randrange(-100, 100) for _ in range(nvec)], np.int64)
    for oe in (None, -(32 * L - 40)):
        Y, ye = BC.trsv_tile(t, cu(V), cu(ve), out_exp=oe)
        Yr, yer = LR.trsv_tile(tr, V, ve, out_exp=oe)
        _eq(Y.cpu().numpy(), Yr, f"trsv limbs out_exp={oe}")
        _eq(ye.cpu().numpy(), yer, f"trsv exps out_exp={oe}")
