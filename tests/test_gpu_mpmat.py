"""GPU fixed-point matrix operations (sdpvibe/mpmat.py:108-156) and MPF reductions
(sdpvibe/mpf.py:354-386) against closed forms over Python integers and mpmath."""
import random

import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd import mpmat
from tests.intutil import EZ, bitlen_tc, rnd_half_up, wrap

pytestmark = pytest.mark.gpu


def _mat(rng, R, C, bits):
    return [[rng.randrange(-(1 << bits), 1 << bits) for _ in range(C)] for _ in range(R)]


def _vals(A):
    rows, ex = mpmat.to_ints(A)
    return rows, [int(x) for x in ex]


def _round_to(v, e_from, e_to, L):
    """v * 2^e_from re-rounded (half up) to exponent e_to, wrapped to L limbs"""
    return wrap(rnd_half_up(v, e_to - e_from), L)


@pytest.mark.parametrize("L", [2, 8, 24])
def test_rescale_neg(L):
    rng = random.Random(L)
    R, C = 5, 7
    rows = _mat(rng, R, C, 32 * L - 9)
    A = mpmat.MP(mpmat.from_ints(rows, 0, L).limbs, torch.tensor([rng.randrange(-50, 50) for _ in range(C)],
                                                                    dtype=torch.int64, device="cuda"))
    _, ea = _vals(A)
    ne = [e + rng.randrange(-20, 40) for e in ea]
    B = mpmat.rescale(A, torch.tensor(ne, dtype=torch.int64, device="cuda"), L + 1)
    got, ge = _vals(B)
    assert ge == ne
    for i in range(R):
        for j in range(C):
            assert got[i][j] == _round_to(rows[i][j], ea[j], ne[j], L + 1)
    Nn = mpmat.neg(A)
    gn, _ = _vals(Nn)
    assert all(gn[i][j] == wrap(-rows[i][j], L) for i in range(R) for j in range(C))


@pytest.mark.parametrize("L", [2, 24])
@pytest.mark.parametrize("mode", ["exact", "align", "col", "mat", None, "fixed"])
def test_add_modes(L, mode):
    rng = random.Random(L * 7 + len(str(mode)))
    R, C = 4, 6
    a = _mat(rng, R, C, 32 * L - 6)
    b = _mat(rng, R, C, 32 * L - 6)
    if mode == "exact":
        A, B = mpmat.from_ints(a, -30, L), mpmat.from_ints(b, -30, L)
        S = mpmat.add(A, B, -1, out_exp=-30)
        got, _ = _vals(S)
        assert all(got[i][j] == wrap(a[i][j] - b[i][j], L) for i in range(R) for j in range(C))
        return
    ea = [rng.randrange(-100, -60) for _ in range(C)]
    eb = [rng.randrange(-100, -60) for _ in range(C)]
    A = mpmat.MP(mpmat.from_ints(a, 0, L).limbs, torch.tensor(ea, dtype=torch.int64, device="cuda"))
    B = mpmat.MP(mpmat.from_ints(b, 0, L).limbs, torch.tensor(eb, dtype=torch.int64, device="cuda"))
    out = {"fixed": -70}.get(mode, mode)
    S = mpmat.add(A, B, 1, out_exp=out)
    got, ge = _vals(S)
    if mode == "align":
        for j in range(C):
            e = max(ea[j], eb[j]) + 1
            assert ge[j] == e
            for i in range(R):
                want = wrap(_round_to(a[i][j], ea[j], e, L) + _round_to(b[i][j], eb[j], e, L), L)
                assert got[i][j] == want
        return
    # general: exact sum at min exponent, then one rounding to the output exponent
    ex = [min(ea[j], eb[j]) for j in range(C)]
    W = [[(a[i][j] << (ea[j] - ex[j])) + (b[i][j] << (eb[j] - ex[j])) for j in range(C)] for i in range(R)]
    hb = 32 * L - 32
    if mode == "fixed":
        oe = [-70] * C
    else:
        tops = [[bitlen_tc(W[i][j]) + ex[j] for j in range(C)] for i in range(R)]
        if mode == "col":
            oe = [max(tops[i][j] for i in range(R)) - hb for j in range(C)]
        else:
            t = max(max(r) for r in tops)
            oe = [t - hb] * C
    assert ge == oe
    for i in range(R):
        for j in range(C):
            assert got[i][j] == _round_to(W[i][j], ex[j], oe[j], L)


@pytest.mark.parametrize("L", [3, 24])
def test_mpf_reductions(L):
    """total: entries rounded (half up) to 2^(max e - 32) in L+2 limbs, then summed exactly
    (mpf.py:354-366); max_abs exact (mpf.py:372-381); dot = total(mul); trace"""
    rng = random.Random(L)
    n = 37
    NB = 32 * L - 2
    ints = [rng.randrange(1 << (NB - 1), 1 << NB) * rng.choice((-1, 1)) for _ in range(n * n)]
    exps = [rng.randrange(-60, 20) - NB for _ in ints]
    ints[3] = 0
    exps[3] = EZ
    A = F.from_ints(ints, exps, L, (n, n))
    mp.mp.prec = 64 * L + 100
    t = max(e for e in exps if e != EZ) - 32
    want = sum(rnd_half_up(v, t - e) for v, e in zip(ints, exps) if v)
    got = F.total(A)
    assert got == mp.mpf(want) * mp.mpf(2) ** t
    mx = max(abs(mp.mpf(v) * mp.mpf(2) ** e) for v, e in zip(ints, exps) if v)
    assert F.max_abs(A) == mx
    tr_i = [(ints[i * n + i], exps[i * n + i]) for i in range(n)]
    tt = max(e for _, e in tr_i if _ != 0) - 32
    assert F.trace(A) == mp.mpf(sum(rnd_half_up(v, tt - e) for v, e in tr_i if v)) * mp.mpf(2) ** tt
    d = F.dot(A, A)
    exact = sum((mp.mpf(v) * mp.mpf(2) ** e) ** 2 for v, e in zip(ints, exps) if v)
    assert abs(d - exact) <= exact * mp.mpf(2) ** (4 - NB)


@pytest.mark.parametrize("L", [3, 24])
def test_max_abs_near_ties_any_order(L):
    """max_abs is the exact maximum whatever the order of the entries, also among values that
    float64 cannot tell apart (equal to 2^-60 relative), duplicates and opposite signs"""
    NB = 32 * L - 2
    base = (1 << (NB - 1)) + 12345
    cands = [(base, 0), (-(base + 1), 0), (base - 1, 0), (base + 1, 0), (base, 0), (-(base + 1), 0), (base >> 3, 3)]
    rng = random.Random(5)
    fill = [(rng.randrange(1, 1 << (NB - 8)), rng.randrange(-40, -1)) for _ in range(50)]
    mp.mp.prec = 64 * L + 100
    want = mp.mpf(base + 1)
    for seed in range(4):
        entries = cands + fill
        random.Random(seed).shuffle(entries)
        A = F.from_ints([v for v, _ in entries], [e for _, e in entries], L, (len(entries),))
        assert F.max_abs(A) == want
