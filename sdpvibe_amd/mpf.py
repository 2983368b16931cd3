"""Multiprecision floating point arrays on the GPU: one exponent per entry (mirror of
sdpvibe/mpf.py).

MPF: m int32 (L, *shape) limb planes (little endian, two's complement), e int64 (*shape);
value = m * 2^e. Normalised: |m| has exactly NB = 32L-2 bits (NB+1 after a rounding carry;
zero: m = 0, e = EZ).

Elementwise ops are gfx950 kernels (one thread per entry); matrix products go through the exact
RNS GEMM on INT8 matrix cores with operands aligned per row (A) / per column (B) -- the usual
normwise floating-point error -- and every output entry rounded once.
"""
import os

import mpmath as mp
import numpy as np
import torch

from ._lib import call, ptr, stream
from .limbops import bitlen_words, shift_round
from .mpmat import RES_BUDGET, ctx, residue_gemm
from .rns import KMAX_LAUNCH, dev

EZ = -(1 << 40)                                      # exponent of zero
LMAX_WIDE = 1 << 20                                  # normalize() takes any input width


class MPF:
    """multiprecision floats: m (L, *shape) int32 limb planes, e (*shape) int64"""
    __slots__ = ("m", "e")

    def __init__(self, m, e):
        self.m, self.e = m, e

    @property
    def L(self):
        return self.m.shape[0]

    @property
    def shape(self):
        return tuple(self.e.shape)

    def __getitem__(self, idx):
        idx = idx if isinstance(idx, tuple) else (idx,)
        return MPF(self.m[(slice(None), *idx)].contiguous(), self.e[idx].contiguous())

    @property
    def T(self):
        """transpose of the last two dimensions (a contiguous copy)"""
        return MPF(self.m.transpose(-1, -2).contiguous(), self.e.transpose(-1, -2).contiguous())

    def reshape(self, *shape):
        return MPF(self.m.reshape(self.L, *shape), self.e.reshape(*shape))

    def clone(self):
        return MPF(self.m.clone(), self.e.clone())

    def __neg__(self):
        return add(zeros(self.shape, self.L), self, -1)

    def __add__(self, other):
        return add(self, other, 1)

    def __sub__(self, other):
        return add(self, other, -1)

    def shift(self, k):
        """exact multiplication by 2^k (k an int or a tensor broadcastable to the shape); zero stays zero"""
        return MPF(self.m, torch.where(self.e == EZ, self.e, self.e + k))

    def to_mp(self):
        """numpy object array of mpmath values (host; exact)"""
        L = self.L
        words = self.m.reshape(L, -1).t().contiguous().cpu().numpy()
        exps = self.e.reshape(-1).cpu().tolist()
        vals = []
        for row, ex in zip(words, exps):
            v = int.from_bytes(row.tobytes(), "little", signed=True)
            vals.append(mp.mpf((v, ex)) if v else mp.mpf(0))
        out = np.empty(len(vals), dtype=object)
        out[:] = vals
        return out.reshape(self.shape)

    def to_float(self):
        """float64 approximations (host numpy): the top two limbs times 2^(e + 32(L - 2)), the exponent
        clamped to [-1100, 1100], each power of two built exactly from its bit pattern"""
        hi = self.m[-1].double()
        top = hi * 2.0 ** 32 + (self.m[-2].double() % 2.0 ** 32) if self.L > 1 else hi
        scale = _pow2((self.e + 32 * (self.L - 2)).clamp(-1100, 1100))
        return torch.where(self.e == EZ, torch.zeros_like(top), top * scale).cpu().numpy()


def _pow2(k):
    """exact float64 2^k for an int64 tensor k: normal range from the exponent field, subnormals
    from the mantissa field, 0 below 2^-1074 and inf above 2^1023"""
    normal = ((k.clamp(-1022, 1023) + 1023) << 52).view(torch.float64)
    sub = (torch.ones_like(k) << (k.clamp(-1074, -1023) + 1074)).view(torch.float64)
    v = torch.where(k >= -1022, normal, sub)
    v = torch.where(k < -1074, torch.zeros_like(v), v)
    return torch.where(k > 1023, torch.full_like(v, float("inf")), v)


def NBITS(L):
    """mantissa bits of a normalised entry"""
    return 32 * L - 2


def zeros(shape, L):
    shape = tuple(shape)
    return MPF(torch.zeros((L, *shape), dtype=torch.int32, device=dev),
               torch.full(shape, EZ, dtype=torch.int64, device=dev))


# ------------------------------------------------------------------ host -> device constructors
def _pack(ints, L):
    """Python ints -> (L, n) int32 limb planes (two's complement), via one byte buffer"""
    nbytes = 4 * L
    raw = b"".join(v.to_bytes(nbytes, "little", signed=True) for v in ints)
    return np.frombuffer(raw, dtype=np.int32).reshape(len(ints), L).copy()


def from_ints(ints, exps, L, shape):
    """mantissas (Python ints) and exponents -> normalised MPF of the given shape"""
    words = torch.from_numpy(np.ascontiguousarray(_pack(ints, L).T)).to(dev)
    return normalize(words, torch.tensor(exps, dtype=torch.int64, device=dev), L).reshape(*shape)


def _round_mag(v, sh):
    """|v| / 2^sh rounded half up (sh > 0), sign kept"""
    r = (abs(v) + (1 << (sh - 1))) >> sh
    return -r if v < 0 else r


def _dec_to_bin(t, NBb):
    """decimal string -> (mantissa of at most NBb bits, binary exponent), the reference's
    conversion (sdpvibe/mpf.py:238-261): integers that fit are exact; larger integers are rounded
    once (half up) to NBb bits; fractions are first rounded to the nearest multiple of 2^-sh with
    sh large enough for NBb + 1 significant bits, then halved (half up) until they fit"""
    t = t.strip()
    neg = t[:1] == "-"
    t = t.lstrip("+-")
    mant_s, _, exp_s = t.replace("E", "e").partition("e")
    e10 = int(exp_s) if exp_s else 0
    ip, _, fp = mant_s.partition(".")
    e10 -= len(fp)
    man = int(ip + fp) if (ip + fp) else 0
    if man == 0:
        return 0, EZ
    sgn = -1 if neg else 1
    if e10 >= 0:
        v = man * 10 ** e10
        over = v.bit_length() - NBb
        if over <= 0:
            return sgn * v, 0
        return sgn * _round_mag(v, over), over
    den = 10 ** -e10
    sh = NBb + den.bit_length() - man.bit_length() + 1
    q = ((man << sh) + den // 2) // den
    ex = -sh
    while q.bit_length() > NBb:
        q = (q + 1) >> 1
        ex += 1
    return sgn * q, ex


def from_strings(strs, L, shape=None):
    """decimal strings -> normalised MPF"""
    pairs = [_dec_to_bin(t, NBITS(L)) for t in strs]
    return from_ints([m_ for m_, _ in pairs], [e_ for _, e_ in pairs], L, shape or (len(pairs),))


def strings_to_host(strs, L):
    """decimal strings -> (limbs (n, L) int32 entry-major, exponents (n,) int64) host arrays"""
    pairs = [_dec_to_bin(t, NBITS(L)) for t in strs]
    limbs = _pack([m_ for m_, _ in pairs], L)
    return limbs, np.array([e_ for _, e_ in pairs], dtype=np.int64)


def host_to_mpf(limbs, exps, L, shape):
    """host (limbs (n, L) entry-major, exponents (n,)) -> normalised MPF on the GPU"""
    words = torch.from_numpy(np.ascontiguousarray(np.asarray(limbs).T)).to(dev)
    return normalize(words, torch.from_numpy(np.asarray(exps, dtype=np.int64)).to(dev), L).reshape(*shape)


def from_mp(vals, L, shape=None):
    """mpmath numbers (taken exactly) or anything mp.mpf accepts -> MPF rounded to NB bits (magnitude
    half up)"""
    NBb = NBITS(L)
    ints, exps = [], []
    for v in vals:
        if not isinstance(v, mp.mpf):
            v = mp.mpf(v)
        sign, man, ex, bc = v._mpf_
        if not man:
            ints.append(0)
            exps.append(EZ)
            continue
        man = int(man)
        if bc > NBb:
            man = _round_mag(man, bc - NBb)
            ex += bc - NBb
        ints.append(-man if sign else man)
        exps.append(ex)
    return from_ints(ints, exps, L, shape or (len(ints),))


def normalize(Xw, ein, L):
    """Xw: (LI, n) int32 wide two's complement, ein (n,) -> normalised MPF (n,) (mpf.py:307-313)"""
    LI, n = Xw.shape
    Xw, ein = Xw.contiguous(), ein.contiguous()
    O = torch.empty((L, n), dtype=torch.int32, device=Xw.device)
    eO = torch.empty((n,), dtype=torch.int64, device=Xw.device)
    call("sdpv_mpf_norm", ptr(Xw), ptr(ein), LI, ptr(O), ptr(eO), L, n, stream())
    return MPF(O, eO)


def mul(A, B):
    """entrywise product; B may be a single entry (broadcast) (mpf.py:316-323)"""
    L = A.L
    n = A.e.numel()
    bs = B.e.numel() == 1
    assert bs or B.shape == A.shape
    Am, Ae, Bm, Be = A.m.contiguous(), A.e.contiguous(), B.m.contiguous(), B.e.contiguous()
    O = torch.empty_like(Am)
    eO = torch.empty_like(Ae)
    call("sdpv_mpf_mul", ptr(Am), ptr(Ae), ptr(Bm), ptr(Be), ptr(O), ptr(eO), L, n, int(bs), stream())
    return MPF(O, eO)


def add(A, B, sgn=1):
    """A + sgn*B, normalised (mpf.py:326-331)"""
    assert A.shape == B.shape, (A.shape, B.shape)
    L = A.L
    n = A.e.numel()
    Am, Ae, Bm, Be = A.m.contiguous(), A.e.contiguous(), B.m.contiguous(), B.e.contiguous()
    O = torch.empty_like(Am)
    eO = torch.empty_like(Ae)
    call("sdpv_mpf_add", ptr(Am), ptr(Ae), ptr(Bm), ptr(Be), ptr(O), ptr(eO), L, n, int(sgn), stream())
    return MPF(O, eO)


def scalar(v, L):
    """one-entry MPF of an mpmath-compatible value"""
    return from_mp([v], L, (1,))


def eye(n, L, val=1):
    """val times the n x n identity"""
    d = scalar(val, L)
    A = zeros((n, n), L)
    i = torch.arange(n, device=dev)
    A.m[:, i, i] = d.m
    A.e[i, i] = d.e
    return A


# ------------------------------------------------------------------ reductions (results on the host)
def _top(A):
    """per entry e + bitlen(m) (bitlen of ~m for m < 0); EZ for zeros"""
    m, e = A.m.contiguous(), A.e.contiguous()
    out = torch.empty_like(e)
    call("sdpv_mpf_top", ptr(m), ptr(e), A.L, e.numel(), ptr(out), stream())
    return out


def total(A):
    """sum of all entries as an mpmath number: every entry rounded (shift_round) to a common
    exponent 32 bits below the largest entry exponent, in L + 2 limbs, and the limb columns summed
    exactly (error below n 2^-NB of the largest entry)"""
    if A.e.numel() == 0:
        return mp.mpf(0)
    base = int(A.e.max()) - 32
    if base < EZ // 2:
        return mp.mpf(0)
    L = A.L
    words = shift_round(A.m.reshape(L, -1), base - A.e.reshape(-1), L + 2).to(torch.int64)
    low = (words[:-1] & 0xFFFFFFFF).sum(1).tolist()              # unsigned lower limbs
    high = int(words[-1].sum())                                   # signed top limb
    v = high << (32 * (L + 1))
    for k, s in enumerate(low):
        v += int(s) << (32 * k)
    return mp.mpf((v, base)) if v else mp.mpf(0)


def dot(A, B):
    return total(mul(A, B))


def max_abs(A):
    """largest |entry| as an exact mpmath number, the same for any order or grouping of the entries:
    the candidates within one binade of the largest top are compared in float64; the ones float64
    cannot separate from the largest (within 2^-40) are deduplicated by their exact words and
    compared exactly. (The reference, sdpvibe/mpf.py:370-380, takes the float64 argmax, which picks
    among such near-ties by position.)"""
    if A.e.numel() == 0:
        return mp.mpf(0)
    tops = _top(A).reshape(-1)
    tmax = int(tops.max())
    if tmax <= EZ // 2:
        return mp.mpf(0)
    near = torch.nonzero(tops >= tmax - 1).reshape(-1)
    flat = A.reshape(-1)
    approx = np.abs(MPF(flat.m[:, near], flat.e[near]).shift(-tmax).to_float())
    close = torch.from_numpy(np.nonzero(approx >= approx.max() * (1 - 2.0 ** -40))[0]).to(near.device)
    idx = near[close]
    if idx.numel() > 1:
        key = torch.cat([flat.m[:, idx].t().to(torch.int64), flat.e[idx][:, None]], 1)
        _, first = np.unique(key.cpu().numpy(), axis=0, return_index=True)
        idx = idx[torch.from_numpy(np.sort(first)).to(idx.device)]
    vals = MPF(flat.m[:, idx], flat.e[idx]).to_mp()
    return max(abs(v) for v in vals)


def trace(A):
    i = torch.arange(A.shape[0], device=A.e.device)
    return total(MPF(A.m[:, i, i], A.e[i, i]))


# ------------------------------------------------------------------ GEMM
def _row_exp(m, e, L, plane, nb, rows, K, ldb, ldr, ldk):
    Erow = torch.empty((nb * rows,), dtype=torch.int64, device=m.device)
    call("sdpv_row_exp", ptr(m), ptr(e), L, plane, nb, rows, K, ldb, ldr, ldk, ptr(Erow), stream())
    return Erow


def _align_rows(A):
    """A (..., R, K) -> (mantissas (L, ..., R, K) with a common exponent per row, row exps (..., R))
    (mpf.py:390-398)"""
    L = A.L
    R_, K = A.shape[-2], A.shape[-1]
    nb = A.e.numel() // max(R_ * K, 1)
    m, e = A.m.contiguous(), A.e.contiguous()
    E = _row_exp(m, e, L, e.numel(), nb, R_, K, R_ * K, K, 1).view(*A.shape[:-1])
    E = torch.where(E < EZ // 2, torch.zeros_like(E), E)          # all-zero rows: exponent 0
    s = (E[..., None] - A.e).reshape(-1)
    s = torch.where(A.e.reshape(-1) == EZ, torch.full_like(s, 1 << 30), s)
    return shift_round(A.m.reshape(L, -1), s, L).view(A.m.shape), E


def _geometry(shape, trans):
    """(rows, K, ldb, ldr, ldk) of op(X) rows over a stored (nb, r, c) matrix X; op = X or X^T"""
    _, r, c = shape
    if trans:
        return c, r, r * c, 1, c
    return r, c, r * c, c, 1


def gemm(A, B, transA=False, transB=False):
    """C = op(A) op(B) (2-d, or batched 3-d (nb, ., .)); exact products, rounded once per entry
    (operands aligned per row of op(A) / column of op(B): normwise floating-point error)
    (mpf.py:401-452).  Row alignment and the residue split are fused into one kernel per operand;
    the product and the CRT + normalisation are one kernel each."""
    if A.e.dim() == 2:
        A1 = MPF(A.m[:, None], A.e[None])
        B1 = MPF(B.m[:, None], B.e[None])
        C = gemm(A1, B1, transA, transB)
        return MPF(C.m[:, 0], C.e[0])
    L = A.L
    assert B.L == L
    Am, Ae, Bm, Be = A.m.contiguous(), A.e.contiguous(), B.m.contiguous(), B.e.contiguous()
    nb = Ae.shape[0]
    assert Be.shape[0] == nb
    M, K, aldb, aldr, aldk = _geometry(Ae.shape, transA)
    N, K2, bldb, bldr, bldk = _geometry(Be.shape, not transB)   # rows of op(B)^T = columns of op(B)
    assert K == K2, (A.shape, B.shape, transA, transB)
    R = ctx(L)
    EA = _row_exp(Am, Ae, L, Ae.numel(), nb, M, K, aldb, aldr, aldk)
    EB = _row_exp(Bm, Be, L, Be.numel(), nb, N, K, bldb, bldr, bldk)
    out_m = torch.empty((L, nb, M, N), dtype=torch.int32, device=dev)
    out_e = torch.empty((nb, M, N), dtype=torch.int64, device=dev)
    if nb * M * N == 0:
        return MPF(out_m, out_e)

    def fa(k0, kc, kp, j0, jw):
        Rh = torch.empty((R.n, kp // 16, nb * M, 16), dtype=torch.int8, device=dev)
        Rl = torch.empty_like(Rh)
        R.split(Am, Ae, EA, L, Ae.numel(), nb, M, aldb, aldr, aldk, k0, kc, kp, Rh, Rl)
        return Rh, Rl

    def fb(k0, kc, kp, j0, jw):
        Rh = torch.empty((R.n, kp // 16, nb * jw, 16), dtype=torch.int8, device=dev)
        Rl = torch.empty_like(Rh)
        if (j0, jw) == (0, N):
            R.split(Bm, Be, EB, L, Be.numel(), nb, N, bldb, bldr, bldk, k0, kc, kp, Rh, Rl)
        else:                                  # column panel: rows j0..j0+jw of op(B)^T, read in place
            Eb = EB.view(nb, N)[:, j0:j0 + jw].contiguous()
            R.split(Bm, Be, Eb, L, Be.numel(), nb, jw, bldb, bldr, bldk, k0, kc, kp, Rh, Rl, off=j0 * bldr)
        return Rh, Rl

    for j0, jw, acc in residue_gemm(R, fa, fb, nb, M, N, K):
        R.crt(acc, nb * M, jw, 0, jw, 1, outm=out_m, oute=out_e, L=L, plane=nb * M * N, ldo=N, jo=j0, Mb=M,
              erow=EA, ecol=EB, scaled=True)
    return MPF(out_m, out_e)


def col_exps(A):
    """row exponents of A^T for a 2-d MPF A (K, M): per column, the largest entry top less NB (EZ
    for an all-zero column), as gemm computes them for op(A) = A^T"""
    K, M = A.shape
    if K == 0:
        return torch.full((M,), EZ, dtype=torch.int64, device=A.e.device)
    return _row_exp(A.m.contiguous(), A.e.contiguous(), A.L, K * M, 1, M, K, K * M, 1, M)


class KSum:
    """C = sum over segments s of A_s^T B_s (A_s (K_s, M), B_s (K_s, N), 2-d MPF), equal bit for bit
    to gemm(cat(A_s), cat(B_s), transA=True) over the segments stacked along K: every entry is
    rounded at the row exponents of the whole stack (EA = max over segments of col_exps(A_s), EB
    likewise), the exact products are summed as residues over all segments, and one reconstruction
    rounds the sum. The segments need not be on the device together (a segment may be fetched,
    added and dropped), and the result does not depend on how the rows are split."""

    def __init__(self, EA, EB, L):
        self.EA, self.EB, self.L = EA.contiguous(), EB.contiguous(), L
        self.M, self.N = EA.numel(), EB.numel()
        self.R = ctx(L)
        self.K = 0
        self.acc = torch.zeros((self.R.n, self.M, self.N), dtype=torch.int16, device=dev)

    def add(self, A, B):
        R, L, M, N = self.R, self.L, self.M, self.N
        K = A.shape[0]
        if A.shape != (K, M) or B.shape != (K, N) or A.L != L or B.L != L:
            raise ValueError(f"KSum.add: segment shapes {A.shape}, {B.shape} do not match ({M}, {N}) at L={L}")
        if self.K + K > R.kmax:
            raise ValueError(f"KSum: inner dimension {self.K + K} exceeds the RNS bound {R.kmax}")
        Am, Ae, Bm, Be = A.m.contiguous(), A.e.contiguous(), B.m.contiguous(), B.e.contiguous()
        per_k = 2 * R.n * (M + N) * 2
        kc = max(64, min(KMAX_LAUNCH, RES_BUDGET // max(per_k, 1) // 64 * 64))
        for k0 in range(0, K, kc):
            kk = min(kc, K - k0)
            kp = R.kp(kk)
            Ah = torch.empty((R.n, kp // 16, M, 16), dtype=torch.int8, device=dev)
            Al = torch.empty_like(Ah)
            Bh = torch.empty((R.n, kp // 16, N, 16), dtype=torch.int8, device=dev)
            Bl = torch.empty_like(Bh)
            # rows of the splits = columns of A_s and B_s (stride 1), inner index = their rows
            R.split(Am, Ae, self.EA, L, K * M, 1, M, K * M, 1, M, k0, kk, kp, Ah, Al)
            R.split(Bm, Be, self.EB, L, K * N, 1, N, K * N, 1, N, k0, kk, kp, Bh, Bl)
            R.product(Ah, Al, Bh, Bl, self.acc, 1, M, N, M * N, accumulate=True, scaled=True)
            del Ah, Al, Bh, Bl
        self.K += K

    def result(self):
        """C (M, N), normalised MPF"""
        L, M, N = self.L, self.M, self.N
        out_m = torch.empty((L, 1, M, N), dtype=torch.int32, device=dev)
        out_e = torch.empty((1, M, N), dtype=torch.int64, device=dev)
        self.R.crt(self.acc, M, N, 0, N, 1, outm=out_m, oute=out_e, L=L, plane=M * N, ldo=N, jo=0, Mb=M,
                   erow=self.EA, ecol=self.EB, scaled=True)
        return MPF(out_m[:, 0], out_e[0])


def symmetrize(A):
    """(A + A^T)/2"""
    return add(A, A.T).shift(-1)


# ------------------------------------------------------------------ Cholesky (equilibrated, batched)
def fx_exp(L):
    """exponent of the equilibrated fixed-point format: |x| <= 1 with 40 bits of integer headroom
    below the top limb (linalg.py:19)"""
    return -(32 * L - 40)


_ONES = {}


def _fixed_one(L, device):
    """limbs of 1 at the fixed-point exponent fx_exp(L), cached per (L, device) so that later calls
    copy nothing from the host (and can be captured in a HIP graph)"""
    key = (L, str(device))
    if key not in _ONES:
        words = (1 << (32 * L - 40)).to_bytes(4 * L, "little", signed=True)
        _ONES[key] = torch.frombuffer(bytearray(words), dtype=torch.int32).to(device)
    return _ONES[key]


class Chol:
    """A = L L^T for one (n, n) or a batch (nb, n, n) of SPD matrices, with L = D^-1 Lt, D = diag(2^-k):
    Lt fixed point at fx_exp(L), k one integer per row (the reference's mpf.Chol, sdpvibe/mpf.py:461-532:
    k_i = floor((top(A_ii) + 1) / 2), D A D rounded to the fixed-point exponent, identity padding to
    a multiple of bt, then the batched recursive Cholesky). Attributes read by callers: Lt, k, n, L,
    nb, npad, tiles.

    dims: optional (nb,) true dimensions of blocks embedded in the top-left corner of A (zeros
    elsewhere); rows beyond a block's dimension get the identity padding, so each block factors
    exactly as it would alone. check=False leaves the positive-definiteness test on the device
    (`not_pd`, a bool tensor) instead of raising, so that no host round trip is needed."""

    SOLVE_REFINE = int(os.environ.get("SDPB_GPU_SOLVE_REFINE", "0"))

    def __init__(self, A, bt=32, refine=1, dims=None, check=True):
        from .batched import BMP, bchol
        self.two_d = A.e.dim() == 2
        if self.two_d:
            A = A.reshape(1, *A.shape)
        nb, n, _ = A.shape
        L = A.L
        dv = A.e.device
        self.nb, self.n, self.L, self.bt, self.refine = nb, n, L, bt, refine
        self.npad = npad = max(bt, -(-n // bt) * bt)
        r = torch.arange(n, device=dv)
        dtop = _top(A)[:, r, r]                                       # (nb, n)
        live = torch.ones((nb, n), dtype=torch.bool, device=dv) if dims is None else \
            r[None, :] < torch.as_tensor(dims, device=dv)[:, None]
        self.not_pd = (((A.m[-1][:, r, r] < 0) | (dtop <= EZ // 2)) & live).any()
        if check and bool(self.not_pd):
            raise ValueError("matrix is not positive definite (diagonal)")
        self.k = k = torch.where(live, torch.div(dtop + 1, 2, rounding_mode="floor"), torch.zeros_like(dtop))
        efx = fx_exp(L)
        # (D A D)_ij = m_ij 2^(e_ij - k_i - k_j) at the fixed-point exponent
        shift = efx - A.e + k[:, :, None] + k[:, None, :]
        shift = torch.where(A.e == EZ, torch.full_like(shift, 1 << 30), shift)
        P = torch.zeros((L, nb, npad, npad), dtype=torch.int32, device=dv)
        P[:, :, :n, :n] = shift_round(A.m.reshape(L, -1), shift.reshape(-1), L).view(L, nb, n, n)
        # identity on the padding rows: beyond n, and beyond each block's own dimension
        one = _fixed_one(L, dv)
        pad = torch.ones((nb, npad), dtype=torch.bool, device=dv)
        pad[:, :n] = ~live
        diag = P.diagonal(dim1=-2, dim2=-1)                           # (L, nb, npad) view
        diag.copy_(torch.where(pad[None], one[:, None, None], diag))
        self.Lt, self.tiles = bchol(BMP(P, efx), bt, refine)

    def _to_bmp(self, B):
        """right-hand sides (nb, n, m) -> fixed point with one exponent per (block, column), zero rows
        below n: each column rounded to its largest entry's exponent (the rows of B^T aligned, as
        _align_rows(B.T) does, read in place instead of through two transposed copies)"""
        from .batched import BMP
        nb, n, m = B.shape
        L = self.L
        bm, be = B.m.contiguous(), B.e.contiguous()
        E = _row_exp(bm, be, L, be.numel(), nb, m, n, n * m, 1, m).view(nb, m)
        E = torch.where(E < EZ // 2, torch.zeros_like(E), E)          # all-zero columns: exponent 0
        s = (E[:, None, :] - be).reshape(-1)
        s = torch.where(be.reshape(-1) == EZ, torch.full_like(s, 1 << 30), s)
        aligned = shift_round(bm.reshape(L, -1), s, L).view(L, nb, n, m)
        if self.npad == n:
            return BMP(aligned, E)
        P = torch.zeros((L, nb, self.npad, m), dtype=torch.int32, device=bm.device)
        P[:, :, :n] = aligned
        return BMP(P, E)

    def _from_bmp(self, P):
        """fixed-point solution (first n rows) -> normalised MPF (nb, n, m)"""
        L, m = self.L, P.shape[2]
        words = P.limbs[:, :, :self.n].reshape(L, -1).contiguous()
        ein = P.expt().unsqueeze(1).expand(self.nb, self.n, m).reshape(-1).contiguous()
        return normalize(words, ein, L).reshape(self.nb, self.n, m)

    def _batched(self, B, f):
        if not self.two_d:
            return f(B)
        out = f(B.reshape(1, *B.shape))
        return out.reshape(*out.shape[1:])

    def solve_lower(self, B):
        """L^-1 B = Lt^-1 (D B)"""
        from .batched import btrsm_left
        return self._batched(B, lambda X: self._from_bmp(btrsm_left(
            self.Lt, self._to_bmp(X.shift(-self.k[:, :, None])), self.tiles, self.bt, self.SOLVE_REFINE)))

    def solve_upper(self, B):
        """L^-T B = D (Lt^-T B)"""
        from .batched import btrsm_left_T
        return self._batched(B, lambda X: self._from_bmp(btrsm_left_T(
            self.Lt, self._to_bmp(X), self.tiles, self.bt, self.SOLVE_REFINE)).shift(-self.k[:, :, None]))

    def solve(self, B):
        """A^-1 B"""
        return self.solve_upper(self.solve_lower(B))


class BucketedChol:
    """Cholesky factors and solves of SPD blocks of uneven dimension. Blocks whose dimension
    rounds up to the same multiple of bt share one batched factorisation: each is embedded with
    the identity padding Chol(block) uses on its own (the `dims` argument of Chol), so every factor
    and every solve is bitwise Chol(block)'s, with one launch sequence per size class instead of
    one per block. (The reference batches equal dimensions with inert padding, solver.py:10-12.)"""

    def __init__(self, blocks, bt=32, refine=1):
        if not blocks:
            raise ValueError("no blocks")
        self.dims = [int(b.shape[-1]) for b in blocks]
        self.L, self.bt = blocks[0].L, bt
        for b, n in zip(blocks, self.dims):
            if b.e.dim() != 2 or b.shape[0] != n or b.L != self.L:
                raise ValueError("blocks must be square 2-d MPF of one limb count")
        classes = {}
        for i, n in enumerate(self.dims):
            classes.setdefault(max(bt, -(-n // bt) * bt), []).append(i)
        self.groups = []
        for npad, idx in sorted(classes.items()):
            A = self._embed([blocks[i] for i in idx], npad, npad)
            dims = torch.tensor([self.dims[i] for i in idx], dtype=torch.int64, device=A.e.device)
            self.groups.append((idx, Chol(A, bt, refine, dims=dims)))

    def _embed(self, Bs, rows, cols):
        dev_ = Bs[0].m.device
        m = torch.zeros((self.L, len(Bs), rows, cols), dtype=torch.int32, device=dev_)
        e = torch.full((len(Bs), rows, cols), EZ, dtype=torch.int64, device=dev_)
        for j, B in enumerate(Bs):
            r, c = B.shape
            m[:, j, :r, :c] = B.m
            e[j, :r, :c] = B.e
        return MPF(m, e)

    def _solve(self, Bs, which):
        if len(Bs) != len(self.dims):
            raise ValueError("one right-hand side per block")
        out = [None] * len(Bs)
        for idx, ch in self.groups:
            cols = {int(Bs[i].shape[-1]) for i in idx}
            if len(cols) != 1 or any(Bs[i].shape[0] != self.dims[i] for i in idx):
                raise ValueError("right-hand sides of one size class need the block's rows and one width")
            R = getattr(ch, which)(self._embed([Bs[i] for i in idx], ch.n, cols.pop()))
            for j, i in enumerate(idx):
                n = self.dims[i]
                out[i] = MPF(R.m[:, j, :n].contiguous(), R.e[j, :n].contiguous())
        return out

    def solve_lower(self, Bs):
        """[L_b^-1 B_b] for the list of right-hand sides B_b (n_b, m)"""
        return self._solve(Bs, "solve_lower")

    def solve_upper(self, Bs):
        """[L_b^-T B_b]"""
        return self._solve(Bs, "solve_upper")

    def solve(self, Bs):
        """[A_b^-1 B_b]"""
        return self._solve(Bs, "solve")


# ------------------------------------------------------------------ Gram accumulation (Q = sum_c P_c^T P_c)
class GramAcc:
    """Q = sum over chunks of P_c^T P_c, accumulated exactly as int16 residues of ctx(L + 1) and
    reconstructed once (mpf.py:536-630).

    Each column j is rounded to a running reference exponent E_j (the largest column top seen),
    leaving HEAD bits of headroom in the extra limb; a chunk whose column top exceeds its
    reference by more than HEAD forces a flush: the partial sum is reconstructed and added in MPF.
    Only tiles on or below the diagonal are accumulated (the product's syrk mode); callers mirror
    the lower triangle."""
    HEAD = 28

    def __init__(self, N, L, E_ref=None):
        self.N, self.L = N, L
        self.R = ctx(L + 1)
        self.E = E_ref.clone() if E_ref is not None else torch.full((N,), EZ, dtype=torch.int64, device=dev)
        self.Tmax = torch.full((N,), EZ, dtype=torch.int64, device=dev)
        self.acc = None
        self.part = None
        self.flushes = 0

    def add(self, P):
        """P: MPF (rows, N). The chunk is rounded to the reference exponents and split into residues
        by one fused kernel reading P's columns in place; its residue Gram matrix is added modulo
        every prime (lower tiles)."""
        L, N, R = self.L, self.N, self.R
        rows = P.shape[0]
        tp = _top(P).amax(0) if rows else torch.full((N,), EZ, dtype=torch.int64, device=dev)
        live = tp > EZ // 2
        self.Tmax = torch.maximum(self.Tmax, tp)
        known = self.E > EZ // 2
        if bool((live & known & (tp > self.E + self.HEAD)).any()):
            self._flush()
            self.E = torch.where(live & (tp > self.E), tp, self.E)
            known = self.E > EZ // 2
        self.E = torch.where(live & ~known, tp, self.E)
        if self.acc is None:
            self.acc = torch.zeros((R.n, N, N), dtype=torch.int16, device=dev)
        if rows == 0:
            return
        Pm, Pe = P.m.contiguous(), P.e.contiguous()
        Erow = (self.E - NBITS(L)).contiguous()
        per_k = 2 * R.n * N
        kc = max(64, min(KMAX_LAUNCH, RES_BUDGET // max(per_k, 1) // 64 * 64))
        for k0 in range(0, rows, kc):
            kk = min(kc, rows - k0)
            kp = R.kp(kk)
            Rh = torch.empty((R.n, kp // 16, N, 16), dtype=torch.int8, device=dev)
            Rl = torch.empty_like(Rh)
            # rows of the split = columns j of P (stride 1), inner index = rows of P (stride N)
            R.split(Pm, Pe, Erow, L, rows * N, 1, N, 0, 1, N, k0, kk, kp, Rh, Rl)
            R.product(Rh, Rl, Rh, Rl, self.acc, 1, N, N, N * N, accumulate=True, syrk=True, scaled=True)
            del Rh, Rl

    def _crt(self):
        L, N, R = self.L, self.N, self.R
        out_m = torch.empty((L, N, N), dtype=torch.int32, device=dev)
        out_e = torch.empty((N, N), dtype=torch.int64, device=dev)
        rE = torch.where(self.E > EZ // 2, self.E - NBITS(L), torch.zeros_like(self.E)).contiguous()
        R.crt(self.acc, N, N, 0, N, 1, outm=out_m, oute=out_e, L=L, plane=N * N, ldo=N, jo=0, Mb=N, erow=rE,
              ecol=rE, scaled=True)
        return MPF(out_m, out_e)

    def _flush(self):
        """reconstruct the residue sum into the MPF partial sum and start over"""
        if self.acc is None:
            return
        q = self._crt()
        self.part = q if self.part is None else add(self.part, q)
        self.acc = None
        self.flushes += 1

    def next_reference(self, margin=4):
        """reference exponents for the next accumulation: this one's column maxima plus margin (EZ
        for columns never seen)"""
        seen = self.Tmax > EZ // 2
        return torch.where(seen, self.Tmax + margin, self.Tmax)

    def precision_loss(self):
        """bits between a column's reference exponent and its actual maximum, largest over columns"""
        both = (self.Tmax > EZ // 2) & (self.E > EZ // 2)
        if not bool(both.any()):
            return 0
        return int((self.E - self.Tmax).masked_select(both).max())

    def result(self):
        """Q (N, N), valid on and below the diagonal"""
        if self.acc is None:
            return self.part
        q = self._crt()
        self.acc = None
        return q if self.part is None else add(self.part, q)
