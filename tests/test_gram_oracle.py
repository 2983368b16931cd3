"""The oracle's GramAcc restatement (sdpvibe/mpf.py:536-630) pinned by mathematics: against the exact
Gram matrix of the unrounded chunks (Python integers), the error of every lower entry stays within
the normwise bound of one rounding per operand entry; the flush count follows the HEAD rule
(tests/test_gramacc.py:11-23 of the reference uses the same chunk scales)."""
import random

import numpy as np
import pytest

from oracle import gram_ref as GR
from tests.intutil import EZ, from_planes, rand_mpf_block


def _exact_gram(chunks, L):
    """exact sum_c P_c^T P_c as a dict of (i, j) -> (int, exponent) with a common exponent"""
    N = chunks[0][0].shape[2]
    emin = min(int(e[e != EZ].min()) for _, e in chunks)
    cols = [[] for _ in range(N)]
    for m, e in chunks:
        ints = np.array(from_planes(m.reshape(L, -1)), dtype=object).reshape(e.shape)
        for r in range(e.shape[0]):
            for j in range(N):
                cols[j].append(0 if e[r, j] == EZ else int(ints[r, j]) << (int(e[r, j]) - emin))
    return cols, 2 * emin


@pytest.mark.parametrize("scales,flushes", [([0, 0, 0], 0), ([0, 10, -40], 0), ([0, 60, -5], 1)])
def test_gram_oracle_error_and_flushes(scales, flushes):
    L, N = 3, 12
    rng = random.Random(sum(scales) + 7)
    chunks = [rand_mpf_block(rng, (r, N), L, shift=sc, zero_frac=0.05) for r, sc in zip((30, 7, 13), scales)]
    G = GR.GramAcc(N, L)
    for m, e in chunks:
        G.add(m, e)
    om, oe = G.result()
    assert G.flushes == flushes
    cols, e0 = _exact_gram(chunks, L)
    NB = 32 * L - 2
    vals = np.array(from_planes(om.reshape(L, -1)), dtype=object).reshape(N, N)
    nrm = [sum(v * v for v in c) for c in cols]
    for i in range(N):
        for j in range(i + 1):
            exact = sum(a * b for a, b in zip(cols[i], cols[j]))
            got = 0 if oe[i, j] == EZ else int(vals[i, j])
            # got * 2^oe - exact * 2^e0, exactly, in units of 2^c with c = min(oe, e0)
            oe_ij = int(oe[i, j]) if got else e0
            c = min(oe_ij, e0)
            err = abs((got << (oe_ij - c)) - (exact << (e0 - c)))
            # |err| <= 2^-(NB - 6) * ||P_i|| ||P_j|| (two roundings per product, one final rounding)
            bound2 = (nrm[i] * nrm[j]) << (2 * (e0 - c)) >> (2 * (NB - 6))
            assert err * err <= bound2 + 1, (i, j)


def test_gram_oracle_reference_exponents():
    L, N = 2, 8
    rng = random.Random(3)
    m, e = rand_mpf_block(rng, (20, N), L)
    G = GR.GramAcc(N, L)
    G.add(m, e)
    nr = G.next_reference()
    assert (nr == G.Tmax + 4).all()
    G2 = GR.GramAcc(N, L, E_ref=nr)
    G2.add(m, e)
    assert G2.precision_loss() == 4
    assert G2.flushes == 0
