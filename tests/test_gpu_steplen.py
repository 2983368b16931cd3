"""Step length on the GPU (sdpvibe/solver.py:996-1103) against exact spectra, the known-answer design of
the reference's tests/test_step_length.py: X = V V^T and dM = V diag(lam) V^T for a random integer V,
built exactly with Python integers, so the eigenvalues of L^-1 dM L^-T (X = L L^T) are exactly lam.
Spectra with large positive eigenvalues next to the smallest defeat the float64 copy; the checked path
(gamma given) must still match the exact step length to EIG_TOL. Also: the float64 -> MPF conversion
is exact, MPF -> float64 follows MPF.to_float, and the FP64 eigensolver matches numpy."""
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd import solver as S

pytestmark = pytest.mark.gpu

L = 40                                                               # 1278 bits: room for 2^700 spreads
GAMMA = mp.mpf("0.7")


def mpf_exact(Q, K):
    n = len(Q)
    ints = [int(v) for row in Q for v in row]
    return F.from_ints(ints, [-K if v else F.EZ for v in ints], L, (n, n))


def case(n, lam, seed):
    rng = np.random.default_rng(seed)
    V = rng.integers(-3, 4, size=(n, n)).astype(object)
    while abs(np.linalg.det(V.astype(float))) < 1:
        V = rng.integers(-3, 4, size=(n, n)).astype(object)
    K = max(0, max(-int(np.floor(np.log2(abs(float(x))))) for x in lam if x) + 8)
    den = max(x.denominator for x in lam)
    K = max(K, den.bit_length() - 1)
    D = [int(x * (1 << K)) for x in lam]
    dM = (V * np.array(D, dtype=object)[None, :]).dot(V.T)
    X = V.dot(V.T)
    return mpf_exact(X, 0), mpf_exact(dM, K)


def run(blocks, expect_refined=None):
    mp.mp.prec = 32 * L
    n = len(blocks[0])
    Xs, Ds = zip(*[case(n, lam, seed) for seed, lam in enumerate(blocks)])
    X = F.MPF(torch.stack([x.m for x in Xs], 1), torch.stack([x.e for x in Xs]))
    D = F.MPF(torch.stack([d.m for d in Ds], 1), torch.stack([d.e for d in Ds]))
    ch = F.Chol(X)
    exact = min(mp.mpf(x.numerator) / x.denominator for lam in blocks for x in lam)
    a_ex = S.step_from_eig(exact, GAMMA)
    st = {}
    lck = S.min_eig([ch], [D], gamma=GAMMA, stats=st)
    ack = S.step_from_eig(lck, GAMMA)
    eck = abs(ack / a_ex - 1)
    assert eck <= S.EIG_TOL, (mp.nstr(eck, 3), mp.nstr(lck, 10), mp.nstr(exact, 10))
    if expect_refined is not None:
        assert st.get("eig_refined", 0) == expect_refined, st


def spectrum(n, big, small, lo, seed=0):
    rng = np.random.default_rng(100 + seed)
    lam = [Fraction(int(rng.integers(1, 8))) * (Fraction(2) ** k) for k in big]
    lam += [Fraction(int(rng.integers(1, 25)), 8) for _ in range(small)]
    lam += [lo]
    lam += [Fraction(int(rng.integers(1, 25)), 8) for _ in range(n - len(lam))]
    rng.shuffle(lam)
    return lam


def test_mpf_from_float_exact():
    V = np.random.default_rng(7).standard_normal((3, 5, 5)) * 2.0 ** np.random.default_rng(8).integers(-900, 900,
                                                                                                      (3, 5, 5))
    V[0, 0, 0] = 0
    got = S._mpf_from_float(torch.from_numpy(V).cuda(), L).to_mp().reshape(-1)
    assert all(mp.mpf(float(x)) == y for x, y in zip(V.reshape(-1), got))


def test_to_double_matches_to_float():
    rng = np.random.default_rng(3)
    vals = [mp.mpf(float(x)) * mp.mpf(2) ** int(k) for x, k in zip(rng.standard_normal(60),
                                                                   rng.integers(-1200, 1200, 60))]
    vals[5] = mp.mpf(0)
    A = F.from_mp(vals, 24, (3, 4, 5))
    sh = torch.tensor([0, 7, -900], dtype=torch.int64, device="cuda")
    got = S.to_double(A, sh).cpu().numpy()
    ref = np.stack([A[b].shift(-int(sh[b])).to_float() for b in range(3)])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [1, 2, 7, 40, 97, 256, 300])
def test_symeig_matches_numpy(n):
    """eigenvalues within the stated bound of LAPACK's, orthonormal eigenvectors, Frobenius norm, and
    the residual off-diagonal norm: zero-order small once converged, and part of the bound"""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((3, n, n)) * np.array([1.0, 1e-5, 1e6])[:, None, None]
    w, V, fro, off = S.symeig(torch.from_numpy(A).cuda(), vectors=True)
    As = (A + A.transpose(0, 2, 1)) / 2
    wr = np.linalg.eigvalsh(As)
    w, V, off = w.cpu().numpy(), V.cpu().numpy(), off.cpu().numpy()
    for b in range(3):
        fb = np.sqrt((As[b] ** 2).sum())
        assert off[b] <= 2.0 ** -50 * fb
        bound = (n + 2) * 2.0 ** -52 * fb + off[b]
        assert np.abs(w[b] - wr[b]).max() <= bound
        assert np.abs(As[b] @ V[b] - V[b] * w[b][None, :]).max() <= 4 * bound
        assert np.abs(V[b].T @ V[b] - np.eye(n)).max() <= 1e-12 * n
    assert np.allclose(fro.cpu().numpy(), np.sqrt((As ** 2).sum(axis=(1, 2))), rtol=1e-14)


def test_symeig_size_limit():
    from sdpvibe_amd._lib import HipLibraryError
    with pytest.raises(HipLibraryError, match="unsupported size"):
        S.symeig(torch.zeros((1, 4097, 4097), dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("n", [40, 144])
def test_ordinary_spectrum(n):
    run([spectrum(n, [], n - 1, Fraction(-11, 8))], 0)


@pytest.mark.parametrize("n", [40, 144])
def test_big_positive(n):
    run([spectrum(n, [100] * 5 + [60] * 3, 4, Fraction(-11, 8))], 1)
    run([spectrum(n, [100] * 5, 4, Fraction(-1, 2))])


@pytest.mark.parametrize("n", [40, 144])
def test_multiscale_and_cluster(n):
    run([spectrum(n, [700, 500, 300, 200, 120, 60, 30], 4, Fraction(-5, 2))], 1)
    run([spectrum(n, [100] * 5, 0, Fraction(-11, 8))[:-3]
         + [Fraction(-11, 8) - Fraction(1, 2 ** 40), Fraction(-11, 8) + Fraction(1, 2 ** 40),
            Fraction(-11, 8) + Fraction(1, 2 ** 20)]])


@pytest.mark.parametrize("n", [40, 144])
def test_big_negative_and_batches(n):
    run([spectrum(n, [90] * 3, 4, -Fraction(2) ** 100)], 0)
    run([spectrum(n, [100] * 5, 4, Fraction(-11, 8), 0), spectrum(n, [700, 400, 100], 4, Fraction(-5, 2), 1),
         spectrum(n, [], n - 1, Fraction(-3, 4), 2), spectrum(n, [300] * 2, 4, Fraction(-1, 2), 3)], 3)
    run([spectrum(n, [], n - 1, Fraction(-3, 4), s) for s in range(3)]
        + [spectrum(n, [200] * 4, 4, Fraction(-9, 4), 3)], 1)


def test_chol_log2_cond():
    """diagonal X = diag(4^j): L = diag(2^j), condition (max/min)^2 = 4^(n-1)"""
    n = 6
    ints = [(1 << (2 * i)) if i == j else 0 for i in range(n) for j in range(n)]
    X = F.from_ints(ints, [0 if v else F.EZ for v in ints], 8, (n, n))
    ch = F.Chol(X)
    hi, lo = S.chol_log2_cond_parts(ch)
    assert abs(hi - (n - 1)) < 1e-9 and abs(lo) < 1e-9
    c = S.chol_log2_cond(ch)
    assert abs(float(c[0]) - 2 * (n - 1)) < 1e-9
