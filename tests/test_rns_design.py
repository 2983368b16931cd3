"""CPU checks of the RNS design (tables and identities the HIP kernels rely on), and a numpy model
of the split -> three-product -> CRT pipeline against exact Python integers."""
import random

import numpy as np
import pytest
import torch

from sdpvibe_amd.rns import PMAX, PMIN, RNS, _primes_desc, _sdigits
from tests.intutil import bitlen_tc, to_planes, from_planes


@pytest.fixture(scope="module")
def ctxs():
    return {L: RNS(L, device="cpu") for L in (1, 2, 16, 24, 26, 32, 33, 34)}


def test_prime_counts(ctxs):
    for L, R in ctxs.items():
        assert len(set(R.mods)) == R.n and all(PMIN < m <= PMAX for m in R.mods)
        bound = R.kmax * (1 << (32 * L - 1)) ** 2
        assert R.M > 4 * bound
        assert R.M // R.mods[-1] <= 4 * bound          # minimal prime count
    assert ctxs[24].n == 112


def test_centred_split_ranges():
    """|r| <= (m-1)/2 <= 8128 -> lo in [-64, 63], hi in [-64, 64], hi + lo in int8 (exhaustive)"""
    for m in (PMAX if PMAX % 2 else PMAX - 1, _primes_desc(PMAX, PMIN)[0], _primes_desc(PMAX, PMIN)[-1]):
        r = np.arange(-(m - 1) // 2, (m - 1) // 2 + 1)
        lo = ((r + 64) & 127) - 64
        hi = (r - lo) >> 7
        assert (128 * hi + lo == r).all()
        assert lo.min() >= -64 and lo.max() <= 63 and hi.min() >= -64 and hi.max() <= 64
        s = hi + lo
        assert s.min() >= -128 and s.max() <= 127


def test_three_product_identity(ctxs):
    rng = np.random.default_rng(1)
    R = ctxs[24]
    kc = R.kcoef.numpy()
    for gi in (0, 50, R.n - 1):
        m = R.mods[gi]
        ra = rng.integers(-(m - 1) // 2, (m - 1) // 2 + 1, 1000)
        rb = rng.integers(-(m - 1) // 2, (m - 1) // 2 + 1, 1000)
        la, lb = ((ra + 64) & 127) - 64, ((rb + 64) & 127) - 64
        ha, hb = (ra - la) >> 7, (rb - lb) >> 7
        hh, ll, ss = ha * hb, la * lb, (ha + la) * (hb + lb)
        assert (16256 * hh + 128 * ss - 127 * ll == ra * rb).all()
        assert ((kc[gi, 0] * hh + kc[gi, 1] * ss + kc[gi, 2] * ll - ra * rb) % m == 0).all()


def test_tables(ctxs):
    for L, R in ctxs.items():
        T = 128 * R.Thi.numpy().astype(np.int64) + R.Tlo.numpy()
        assert R.Thi.shape == (R.npad, R.NBp) and (T[R.n:] == 0).all() and (T[:, 4 * L:] == 0).all()
        for gi in (0, R.n // 2, R.n - 1):
            m = R.mods[gi]
            assert all((int(T[gi, j]) - pow(256, j, m)) % m == 0 and abs(int(T[gi, j])) <= m // 2
                       for j in range(4 * L))
            assert int(R.C1[gi]) == (128 * int(T[gi, :4 * L - 1].sum())) % m
            assert (int(R.inv32[gi]) * (R.M // m)) % m == 1
            dig = R.Ct.numpy()[:, gi].astype(np.int64)
            assert sum(int(d) << (8 * k) for k, d in enumerate(dig)) == R.M // m
        md = R.Md.numpy().astype(np.int64)
        assert sum(int(d) << (8 * k) for k, d in enumerate(md)) == R.M
        assert R.npad > R.n and (R.Ct.numpy()[:, R.n] == md).all() and (R.Ct.numpy()[:, R.n + 1:] == 0).all()
        kc, ks = R.kcoef.numpy().astype(np.int64), R.kcoef_crt.numpy().astype(np.int64)
        for gi in (0, R.n - 1):
            m, v = R.mods[gi], int(R.inv32[gi])
            assert [int(x) for x in ks[gi]] == [int(x) * v % m for x in kc[gi]]
        assert R.E % 8 == 0 and 8 * R.E >= R.M.bit_length() + 16


def _model_residues(R, vals, L):
    """numpy model of the split kernel: bytes u - 128 (the top byte signed) . T + C1 (mod m), centred"""
    X = to_planes(vals, L).T.copy().view(np.uint8).reshape(len(vals), 4 * L).astype(np.int64)
    xt = X - 128
    xt[:, -1] = np.where(X[:, -1] >= 128, X[:, -1] - 256, X[:, -1])   # signed top byte
    assert xt.min() >= -128 and xt.max() <= 127
    Th = R.Thi.numpy()[:R.n, :4 * L].astype(np.int64)
    Tl = R.Tlo.numpy()[:R.n, :4 * L].astype(np.int64)
    m = np.array(R.mods, np.int64)
    x = 128 * (xt @ Th.T) + xt @ Tl.T + R.C1.numpy()[None, :]
    r = x % m[None, :]
    return np.where(r > m // 2, r - m, r)                     # (entries, n)


def test_pipeline_model_against_integers(ctxs):
    """split (byte tables) -> three INT8 products per prime -> CRT digits and carry, in numpy int64,
    against the exact Python-integer product"""
    L, M_, K, N_ = 2, 4, 9, 3
    R = ctxs[L]
    rng = random.Random(7)
    lim = 1 << (32 * L - 1)
    A = [[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(M_)]
    B = [[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(N_)]
    A[0][:3] = [lim - 1, -lim, 0]
    B[0][:3] = [-lim, -lim, 5]
    ra = _model_residues(R, [v for r in A for v in r], L).reshape(M_, K, R.n)
    rb = _model_residues(R, [v for r in B for v in r], L).reshape(N_, K, R.n)
    for g, m in enumerate(R.mods):
        assert all(ra[i, k, g] == ((A[i][k] % m) if A[i][k] % m <= m // 2 else A[i][k] % m - m)
                   for i in range(M_) for k in range(K))
    split = lambda r: (((r + 64) & 127) - 64)
    la, lb = split(ra), split(rb)
    ha, hb = (ra - la) >> 7, (rb - lb) >> 7
    HH = np.einsum("ikg,jkg->gij", ha, hb)
    LL = np.einsum("ikg,jkg->gij", la, lb)
    SS = np.einsum("ikg,jkg->gij", ha + la, hb + lb)
    m = np.array(R.mods, np.int64)[:, None, None]
    kc = R.kcoef.numpy().astype(np.int64)
    acc = (kc[:, 0, None, None] * (HH % m) + kc[:, 1, None, None] * (SS % m) + kc[:, 2, None, None] * (LL % m)) % m
    # CRT: y = acc * inv mod m (centred), k = round(sum y/m), Z = digits . y, carry
    y = (acc * R.inv32.numpy().astype(np.int64)[:, None, None]) % m
    y = np.where(y > m // 2, y - m, y)
    # the scaled coefficients give y directly (the product's CRT-ready output)
    ks = R.kcoef_crt.numpy().astype(np.int64)
    ys = (ks[:, 0, None, None] * HH + ks[:, 1, None, None] * SS + ks[:, 2, None, None] * LL) % m
    assert (np.where(ys > m // 2, ys - m, ys) == y).all()
    k = np.floor((y / m).sum(0) + 0.5).astype(np.int64)
    assert (np.abs(k) <= R.n // 2).all()
    # -k in column n of Y against the digits of M in column n of Ct
    Yx = np.concatenate([y, -k[None]], 0)
    Ct = R.Ct.numpy()[:R.E, :R.n + 1].astype(np.int64)
    Z = np.einsum("dg,gij->dij", Ct, Yx)
    for i in range(M_):
        for j in range(N_):
            v = sum(int(Z[d, i, j]) << (8 * d) for d in range(R.E))
            assert v == sum(A[i][kk] * B[j][kk] for kk in range(K))


def test_digit_sums_fit_int32(ctxs):
    """worst case of the CRT digit products: |y| <= 8128 split hi/lo, |Ct| <= 128"""
    for R in ctxs.values():
        assert R.npad * 128 * 64 < 1 << 31 and R.npad * 128 * 64 * 128 < 1 << 31


def test_sdigits_roundtrip():
    rng = random.Random(2)
    for _ in range(50):
        x = rng.randrange(-(1 << 300), 1 << 300)
        d = _sdigits(x, 256, 40)
        assert sum(v << (8 * i) for i, v in enumerate(d)) == x and all(-128 <= v < 128 for v in d)
