"""GPU tiled Cholesky (sdpvibe/blockchol.py) bit for bit against the oracle's restatement (pinned in
test_blockchol_oracle.py), and against the monolithic Chol to working precision."""
import random

import numpy as np
import pytest
import torch

from oracle import blockchol_ref as BCR
from sdpvibe_amd import mpf as F
from sdpvibe_amd.blockchol import BlockChol
from tests.intutil import mpf_host, spd_blocks

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    bad = np.argwhere(a != b)
    assert a.shape == b.shape and bad.size == 0, f"{what}: first mismatches {bad[:5].tolist()}"


def blockchol_case(L, N, b, seed):
    m, e, _ = spd_blocks(1, N, L, seed=seed, uspread=6)
    Q = (m[:, 0], e[0])
    ref = BCR.BlockChol(Q, b)
    bc = BlockChol([F.MPF(cu(Q[0]), cu(Q[1]))], b)
    rng = random.Random(seed)
    bm, be = mpf_host([(rng.randrange(-(1 << 80), 1 << 80), rng.randrange(-20, 20)) for _ in range(N * 3)], L)
    B = (bm.reshape(L, N, 3), be.reshape(N, 3))
    Y = bc.solve_lower(F.MPF(cu(B[0]), cu(B[1])))
    Yr = ref.solve_lower(B)
    _eq(Y.m.cpu().numpy(), Yr[0], "solve_lower m")
    _eq(Y.e.cpu().numpy(), Yr[1], "solve_lower e")
    X = bc.solve_upper(Y)
    Xr = ref.solve_upper(Yr)
    _eq(X.m.cpu().numpy(), Xr[0], "solve_upper m")
    _eq(X.e.cpu().numpy(), Xr[1], "solve_upper e")
    return bc


@pytest.mark.parametrize("L,N,b", [(3, 70, 32), (24, 100, 48), (8, 64, 64)])
def test_blockchol_parity(L, N, b):
    bc = blockchol_case(L, N, b, seed=L + N)
    assert np.isfinite(bc.log2_cond())
