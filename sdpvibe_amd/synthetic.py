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


# ------------------------------------------------------------------ GPU-assembled configurations
# Each generator draws its random numbers on the host (seeded PCG64, so the inputs are the same on
# every machine) and assembles the matrices with this package's exact kernels on the current GPU.

def _gpu(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _mpf(m, e):
    from .mpf import MPF
    return MPF(_gpu(m), _gpu(e))


def powers_of_ten(kmin, kmax, L):
    """MPF (kmax - kmin + 1,) with entry i = 10^(kmin + i) rounded to NB bits (half up)"""
    from .mpf import from_strings
    return from_strings([f"1e{k}" for k in range(kmin, kmax + 1)], L)


def gaussian_pow10(shape, L, seed, vpow):
    """MPF entries g * 10^v on the GPU: g a float64 standard normal (exact), v an integer uniform in
    [-vpow, vpow] per entry, the product rounded once to NB bits. Cheap enough for 10^8 entries."""
    import torch
    from .mpf import MPF, mul
    from .solver import _mpf_from_float
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    g = _mpf_from_float(_gpu(rng.standard_normal(n)), L)
    v = _gpu(rng.integers(0, 2 * vpow + 1, n))
    p10 = powers_of_ten(-vpow, vpow, L)
    D = MPF(p10.m[:, v].contiguous(), p10.e[v].contiguous())
    out = mul(g, D)
    del g, D
    torch.cuda.empty_cache()
    return out.reshape(*shape)


def plain_config(n=512, L=24, seed=0):
    """The plain-kernel configuration's matrices: B (n, n) Gaussian MPF (gaussian_mpf, `seed`) and the
    SPD A = B B^T + n I (exact product rounded once per entry by mpf.gemm, then mpf.add)."""
    from .mpf import add, eye, gemm
    B = _mpf(*gaussian_mpf((n, n), L, seed=seed))
    A = add(gemm(B, B, transB=True), eye(n, L, val=n))
    return B, A


def uneven_spd_blocks(nblocks, L, seed=0, dmin=8, dmax=96, upow=20, rhs=4):
    """SPD blocks A_b = D (B B^T + n I) D of dimension n uniform in [dmin, dmax]: B Gaussian MPF,
    D = diag(10^u), u uniform in [-upow, upow] (condition up to ~10^(4 upow)), each entry of
    B B^T + n I multiplied by 10^(u_i + u_j) rounded to NB bits (one rounding); `rhs` Gaussian
    right-hand sides per block. Returns (dims, blocks, right-hand sides) (lists of 2-d MPF)."""
    from .mpf import MPF, add, eye, gemm, mul
    rng = np.random.default_rng(seed)
    dims = [int(x) for x in rng.integers(dmin, dmax + 1, nblocks)]
    p10 = powers_of_ten(-2 * upow, 2 * upow, L)
    blocks, rhs_l = [], []
    for i, n in enumerate(dims):
        B = _mpf(*gaussian_mpf((n, n), L, seed=10_000 + seed * 7919 + i))
        G = add(gemm(B, B, transB=True), eye(n, L, val=n))
        u = rng.integers(-upow, upow + 1, n)
        idx = _gpu((u[:, None] + u[None, :] + 2 * upow).reshape(-1))
        D = MPF(p10.m[:, idx].reshape(L, n, n).contiguous(), p10.e[idx].reshape(n, n).contiguous())
        blocks.append(mul(G, D))
        rhs_l.append(_mpf(*gaussian_mpf((n, rhs), L, seed=20_000 + seed * 7919 + i)))
    return dims, blocks, rhs_l


def schur_blocks(nblocks, N, L, seed=0, dmin=20, dmax=80, vpow=30):
    """Schur-complement blocks: dimension d uniform in [dmin, dmax], S_b = G G^T + d I (G Gaussian
    MPF), B_b (d, N) with entries g * 10^v (gaussian_pow10). Returns (dims, [S_b], B) with B the
    row-stacked B_b (sum d, N)."""
    from .mpf import add, eye, gemm
    rng = np.random.default_rng(seed)
    dims = [int(x) for x in rng.integers(dmin, dmax + 1, nblocks)]
    S = []
    for i, d in enumerate(dims):
        G = _mpf(*gaussian_mpf((d, d), L, seed=30_000 + seed * 7919 + i))
        S.append(add(gemm(G, G, transB=True), eye(d, L, val=d)))
    B = gaussian_pow10((sum(dims), N), L, seed=40_000 + seed, vpow=vpow)
    return dims, S, B
