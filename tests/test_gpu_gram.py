"""GPU mpf.GramAcc (sdpvibe/mpf.py:536-630) bit for bit against the oracle's restatement (pinned in
test_gram_oracle.py) on the lower triangle, which is the result."""
import random

import numpy as np
import pytest
import torch

from oracle import gram_ref as GR
from sdpvibe_amd import mpf as F
from tests.intutil import rand_mpf_block

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gram_parity(L, N, rows, scales, seed, e_ref=False):
    rng = random.Random(seed)
    chunks = [rand_mpf_block(rng, (r, N), L, shift=sc, zero_frac=0.03) for r, sc in zip(rows, scales)]
    ref = GR.GramAcc(N, L)
    acc = F.GramAcc(N, L)
    if e_ref:
        for m, e in chunks:
            ref.add(m, e)
        nr = ref.next_reference()
        ref = GR.GramAcc(N, L, E_ref=nr)
        acc = F.GramAcc(N, L, E_ref=cu(nr))
    for m, e in chunks:
        ref.add(m, e)
        acc.add(F.MPF(cu(m), cu(e)))
    rm, re_ = ref.result()
    Q = acc.result()
    assert acc.flushes == ref.flushes
    assert acc.precision_loss() == ref.precision_loss()
    assert np.array_equal(acc.next_reference().cpu().numpy(), ref.next_reference())
    low = np.tril(np.ones((N, N), bool))
    qm, qe = Q.m.cpu().numpy(), Q.e.cpu().numpy()
    bad = np.argwhere((qe != re_) & low)
    assert bad.size == 0, f"exponents differ at {bad[:5].tolist()}"
    bad = np.argwhere((qm != rm).any(0) & low)
    assert bad.size == 0, f"mantissas differ at {bad[:5].tolist()}"
    return acc.flushes


@pytest.mark.parametrize("scales,flushes", [([0, 0, 0], 0), ([0, 10, -40], 0), ([0, 60, -5], 1)])
def test_gram_parity_reference_chunks(scales, flushes):
    """the chunk shapes and scale jumps of the reference's tests/test_gramacc.py:11-23"""
    assert gram_parity(18, 96, (300, 17, 129), scales, seed=1) == flushes


@pytest.mark.parametrize("L,N,rows", [(3, 40, (5, 70)), (24, 130, (64, 200, 1)), (32, 64, (300,))])
def test_gram_parity_shapes(L, N, rows):
    gram_parity(L, N, rows, [0] * len(rows), seed=L + N)


def test_gram_reference_exponents():
    gram_parity(8, 50, (40, 40), [0, 3], seed=9, e_ref=True)
