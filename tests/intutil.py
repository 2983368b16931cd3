"""Python-int helpers shared by the tests: limb planes <-> ints and the reference's closed forms.
Independent of both the HIP library and the C oracle (plain Python integers)."""
import random

import numpy as np

EZ = -(1 << 40)
MASK = (1 << 32) - 1


def to_planes(values, L):
    """list of Python ints -> (L, n) int32 limb planes (two's complement, wraps mod 2^(32L))"""
    out = np.zeros((L, len(values)), np.int64)
    for j, v in enumerate(values):
        u = v % (1 << (32 * L))
        for k in range(L):
            out[k, j] = (u >> (32 * k)) & MASK
    out = np.where(out >= 1 << 31, out - (1 << 32), out)
    return out.astype(np.int32)


def from_planes(X):
    """(L, n) int32 limb planes -> list of signed Python ints"""
    X = np.asarray(X).reshape(X.shape[0], -1).astype(np.int64) & MASK
    L = X.shape[0]
    vals = []
    for j in range(X.shape[1]):
        v = 0
        for k in reversed(range(L)):
            v = (v << 32) | int(X[k, j])
        if v >> (32 * L - 1):
            v -= 1 << (32 * L)
        vals.append(v)
    return vals


def bitlen_tc(v):
    """bit length of v, of ~v for negatives (tests/test_rns_kernels.py:78-80)"""
    return (v if v >= 0 else ~v).bit_length()


def rnd_half_up(v, s):
    """round(v / 2^s) half toward +inf, s may be <= 0 (tests/test_gemm.py:195-197)"""
    if s <= 0:
        return v << (-s)
    return (v + (1 << (s - 1))) >> s


def wrap(v, L):
    """two's-complement wrap to L limbs"""
    M = 1 << (32 * L)
    v %= M
    return v - M if v >> (32 * L - 1) else v


def norm_closed_form(v, NB):
    """mpf_norm / mpf_mul rounding of an exact integer v (tests/test_mpf_mul.py:131-139):
    (mantissa, shift); magnitude rounded half up to NB bits; (0, None) for zero"""
    if v == 0:
        return 0, None
    mag = abs(v)
    shift = mag.bit_length() - NB
    r = ((mag + (1 << (shift - 1))) >> shift) if shift > 0 else mag << -shift
    return (r if v > 0 else -r), shift


def rand_mpf_ints(rng, n, L, espread=2000):
    """normalised random MPF entries as (mantissa ints, exponents) incl. edge cases"""
    NB = 32 * L - 2
    edge = [0, 1, -1, (1 << NB) - 1, -((1 << NB) - 1), 1 << (NB - 1), -(1 << (NB - 1)), (1 << (NB - 1)) + 1]
    vals = edge + [rng.randrange(-(1 << NB), 1 << NB) for _ in range(max(0, n - len(edge)))]
    vals = vals[:n]
    exps = [rng.randrange(-espread, espread) for _ in vals]
    return vals, exps


def words_to_int(w):
    """1-d array of uint32/int32 words (little endian, two's complement) -> Python int"""
    w = [int(x) & MASK for x in w]
    v = 0
    for x in reversed(w):
        v = (v << 32) | x
    if w and w[-1] >> 31:
        v -= 1 << (32 * len(w))
    return v


def int_to_words(v, n):
    u = v % (1 << (32 * n))
    return [(u >> (32 * k)) & MASK for k in range(n)]


def spd_tile(n, logcond, seed, F):
    """equilibrated SPD matrix (unit diagonal, cond ~ 10^logcond) as Python ints scaled by 2^F"""
    import mpmath as mpm
    rng = np.random.default_rng(seed)
    Qd, _ = np.linalg.qr(rng.standard_normal((n, n)))
    mpm.mp.prec = 4 * F + 200
    Q = mpm.matrix(Qd.tolist())
    D = mpm.diag([mpm.mpf(10) ** (-mpm.mpf(logcond) * k / max(n - 1, 1)) for k in range(n)])
    A = Q * D * Q.T
    d = [1 / mpm.sqrt(A[i, i]) for i in range(n)]
    return [[int(mpm.nint(A[i, j] * d[i] * d[j] * mpm.mpf(2) ** F)) for j in range(n)] for i in range(n)]


def spd_blocks(nb, n, L, seed, uspread=20, gbits=None):
    """nb SPD matrices A_b = D (G G^T + n I) D with D = diag(2^u), u uniform in [-uspread, uspread],
    as normalised MPF host arrays. Returns (m (L, nb, n, n) int32, e (nb, n, n) int64, exact) with
    exact[b][i][j] = (integer, exponent) of the value before rounding to NB bits."""
    rng = random.Random(seed)
    NB = 32 * L - 2
    gb = gbits or max(8, NB // 2 + 4)
    m = np.zeros((nb, n, n), dtype=object)
    e = np.zeros((nb, n, n), np.int64)
    exact = []
    for b in range(nb):
        G = [[rng.randrange(-(1 << gb), 1 << gb) for _ in range(n + 2)] for _ in range(n)]
        u = [rng.randint(-uspread, uspread) for _ in range(n)]
        ex_b = []
        for i in range(n):
            row = []
            for j in range(n):
                v = sum(G[i][t] * G[j][t] for t in range(n + 2)) + (n << (2 * gb) if i == j else 0)
                ev = u[i] + u[j]
                row.append((v, ev))
                r, sh = norm_closed_form(v, NB)
                m[b, i, j] = r
                e[b, i, j] = EZ if v == 0 else ev + sh
            ex_b.append(row)
        exact.append(ex_b)
    planes = to_planes([int(x) for x in m.reshape(-1)], L).reshape(L, nb, n, n)
    return planes, e, exact


def mpf_host(vals_exps, L):
    """list of (integer, exponent) -> normalised (planes (L, n), exps (n,))"""
    NB = 32 * L - 2
    ms, es = [], []
    for v, ev in vals_exps:
        r, sh = norm_closed_form(v, NB)
        ms.append(r)
        es.append(EZ if v == 0 else ev + sh)
    return to_planes(ms, L), np.array(es, np.int64)


def rand_mpf_block(rng, shape, L, spread=8, shift=0, zero_frac=0.0):
    """normalised random MPF host arrays (planes (L, *shape), exps shape): |mantissa| has NB bits,
    exponent -NB + u + shift with u uniform in [-spread, spread]; a fraction of exact zeros"""
    NB = 32 * L - 2
    n = int(np.prod(shape))
    ms, es = [], []
    for _ in range(n):
        if zero_frac and rng.random() < zero_frac:
            ms.append(0)
            es.append(EZ)
            continue
        v = rng.randrange(1 << (NB - 1), 1 << NB)
        ms.append(-v if rng.random() < 0.5 else v)
        es.append(-NB + rng.randint(-spread, spread) + shift)
    return to_planes(ms, L).reshape((L,) + tuple(shape)), np.array(es, np.int64).reshape(shape)
