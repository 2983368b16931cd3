"""Bucketed Cholesky of uneven blocks (the batched-by-size-class path): every factor and solve is
bitwise the one of Chol(block) on its own (the plain per-block form, itself pinned against the
oracle in test_gpu_chol.py)."""
import random

import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from tests.intutil import mpf_host, spd_blocks

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def bucketed_case(put, L, dims, r, seed):
    blocks, rhs = [], []
    rng = random.Random(seed)
    for i, n in enumerate(dims):
        m, e, _ = spd_blocks(1, n, L, seed=seed + i, uspread=12)
        blocks.append(F.MPF(put(m[:, 0]), put(e[0])))
        vals = [(rng.randrange(-(1 << 120), 1 << 120), rng.randrange(-40, 40)) for _ in range(n * r)]
        bm, be = mpf_host(vals, L)
        rhs.append(F.MPF(put(bm.reshape(L, n, r)), put(be.reshape(n, r))))
    BC = F.BucketedChol(blocks)
    classes = {max(32, -(-n // 32) * 32) for n in dims}
    assert len(BC.groups) == len(classes)
    got = {w: getattr(BC, w)(rhs) for w in ("solve_lower", "solve_upper", "solve")}
    for idx, ch in BC.groups:
        for j, i in enumerate(idx):
            C = F.Chol(blocks[i])
            _same(ch.Lt.limbs[:, j], C.Lt.limbs[:, 0], f"factor of block {i}")
            _same(ch.k[j, :dims[i]], C.k[0], f"scales of block {i}")
            for w, res in got.items():
                ref = getattr(C, w)(rhs[i])
                _same(res[i].m, ref.m, f"{w} block {i} mantissas")
                _same(res[i].e, ref.e, f"{w} block {i} exponents")


@pytest.mark.parametrize("L", [3, 24])
def test_bucketed_matches_per_block(L):
    bucketed_case(lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda(), L,
                  [5, 32, 17, 33, 70, 64, 8, 96, 40], 4, seed=L)


def test_bucketed_rejects_mixed_widths():
    m, e, _ = spd_blocks(2, 6, 2, seed=1)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    blocks = [F.MPF(cu(m[:, b]), cu(e[b])) for b in range(2)]
    BC = F.BucketedChol(blocks)
    z = lambda n, c: F.MPF(torch.zeros((2, n, c), dtype=torch.int32, device="cuda"),
                           torch.full((n, c), F.EZ, dtype=torch.int64, device="cuda"))
    with pytest.raises(ValueError):
        BC.solve_lower([z(6, 2), z(6, 3)])
