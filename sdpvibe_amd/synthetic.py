"""Portable synthetic multiprecision inputs (host numpy, seeded PCG64; identical on every machine).

The reference draws its synthetic matrices from the CUDA torch RNG (sdpvibe/solver.py:401-406),
whose bits depend on the GPU; these generators replace it with a host RNG.
"""
import numpy as np

EZ = -(1 << 40)


def _negate_planes(W, neg):
    """two's-complement negation of the (L, n) uint64-held words where neg is True"""
    L, n = W.shape
    carry = np.ones(n, np.uint64)
    out = W.copy()
    for k in range(L):
        t = ((~W[k]) & np.uint64(0xFFFFFFFF)) + carry
        out[k] = np.where(neg, t & np.uint64(0xFFFFFFFF), W[k])
        carry = t >> np.uint64(32)
    return out


def gaussian_mpf(shape, L, seed=0, espread=16):
    """Normalised MPF entries s * |g| * 2^u: g a full-precision Gaussian (53 leading bits from a
    float64 normal variate, the remaining NB - 53 mantissa bits uniform), s = +-1 uniform, u an
    integer uniform in [-espread, espread]. Returns host (m (L, *shape) int32, e (*shape) int64).
    Requires L >= 2."""
    assert L >= 2
    n = int(np.prod(shape))
    NB = 32 * L - 2
    rng = np.random.default_rng(seed)
    g = np.abs(rng.standard_normal(n))
    sign = rng.integers(0, 2, n).astype(bool)
    u = rng.integers(-espread, espread + 1, n)
    f, e2 = np.frexp(g)                                    # g = f 2^e2, f in [0.5, 1)
    top = (f * 2.0 ** 53).astype(np.uint64)                # 53 bits, bit 52 set
    W = rng.integers(0, 1 << 32, (L, n), dtype=np.uint64)
    # bits [NB-53, NB) hold `top`: offset 9 of word L-2 through bit 29 of word L-1
    lowmask = np.uint64((1 << 9) - 1)
    W[L - 2] = (W[L - 2] & lowmask) | ((top << np.uint64(9)) & np.uint64(0xFFFFFFFF))
    W[L - 1] = top >> np.uint64(23)
    W = _negate_planes(W, sign)
    m = W.astype(np.uint32).view(np.int32).reshape((L,) + tuple(shape))
    e = (e2.astype(np.int64) - NB + u).reshape(shape)
    return np.ascontiguousarray(m), np.ascontiguousarray(e)
