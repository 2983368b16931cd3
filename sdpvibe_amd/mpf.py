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
from .mpmat import ACC_BUDGET, RES_BUDGET, ctx, residue_gemm
from .rns import KMAX_LAUNCH, dev

EZ = -(1 << 40)                                      # exponent of zero
LMAX_WIDE = 1 << 20                                  # normalize() takes any input width


class MPF:
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
        if not isinstance(idx, tuple):
            idx = (idx,)
        return MPF(self.m[(slice(None),) + idx].contiguous(), self.e[idx].contiguous())

    @property
    def T(self):
        """transpose of the last two dimensions"""
        return MPF(self.m.transpose(-1, -2).contiguous(), self.e.transpose(-1, -2).contiguous())

    def reshape(self, *s):
        return MPF(self.m.reshape(self.L, *s), self.e.reshape(*s))

    def clone(self):
        return MPF(self.m.clone(), self.e.clone())

    def __neg__(self):
        return add(zeros(self.shape, self.L), self, -1)

    def __add__(self, o):
        return add(self, o, 1)

    def __sub__(self, o):
        return add(self, o, -1)

    def shift(self, k):
        """multiply by 2^k (k: int or tensor broadcastable to shape) -- exact (mpf.py:194-197)"""
        e = self.e + k
        return MPF(self.m, torch.where(self.e == EZ, EZ, e))

    def to_mp(self):
        """-> numpy object array of mpmath mpf (host; tests / scalars)"""
        L = self.L
        a = self.m.reshape(L, -1).t().contiguous().cpu().numpy().astype(np.uint32)
        ex = self.e.reshape(-1).cpu().numpy()
        out = np.empty(a.shape[0], dtype=object)
        for i in range(a.shape[0]):
            v = int.from_bytes(a[i].tobytes(), "little", signed=True)
            out[i] = mp.mpf(0) if v == 0 else mp.mpf((v, int(ex[i])))
        return out.reshape(self.shape)

    def to_float(self):
        """approximate float64 values: the top two limbs times 2^(e + 32(L-2)), the exponent
        clamped to [-1100, 1100]; host numpy (mpf.py:210-215). The power of two is built exactly
        from its bit pattern (a generic pow may round for subnormal results)."""
        top = self.m[-1].double() * 2.0 ** 32 + (self.m[-2].double() % 2.0 ** 32) if self.L > 1 else self.m[-1].double()
        sc = (self.e + 32 * (self.L - 2)).clamp(-1100, 1100)
        v = torch.where(self.e == EZ, torch.zeros_like(top), top * _pow2(sc))
        return v.cpu().numpy()


def _pow2(k):
    """exact float64 2^k for an int64 tensor k: normal range from the exponent field, subnormals
    from the mantissa field, 0 below 2^-1074 and inf above 2^1023"""
    normal = ((k.clamp(-1022, 1023) + 1023) << 52).view(torch.float64)
    sub = (torch.ones_like(k) << (k.clamp(-1074, -1023) + 1074)).view(torch.float64)
    v = torch.where(k >= -1022, normal, sub)
    v = torch.where(k < -1074, torch.zeros_like(v), v)
    return torch.where(k > 1023, torch.full_like(v, float("inf")), v)


def zeros(shape, L):
    return MPF(torch.zeros((L, *shape), dtype=torch.int32, device=dev),
               torch.full(shape, EZ, dtype=torch.int64, device=dev))


def NBITS(L):
    return 32 * L - 2


def from_ints(ints, exps, L, shape):
    """Python ints (mantissas) and exponents (flat lists) -> normalised MPF (mpf.py:226-235)"""
    n = len(ints)
    nb = 4 * L
    buf = bytearray(nb * n)
    for i, v in enumerate(ints):
        buf[i * nb:(i + 1) * nb] = v.to_bytes(nb, "little", signed=True)
    m = torch.from_numpy(np.frombuffer(bytes(buf), dtype=np.int32).reshape(n, L).T.copy()).to(dev)
    e = torch.tensor(exps, dtype=torch.int64, device=dev)
    return normalize(m, e, L).reshape(*shape)


def _dec_to_bin(t, NBb):
    """decimal string -> (signed mantissa with <= NBb bits, binary exponent) (mpf.py:238-261)"""
    t = t.strip()
    neg = t.startswith("-")
    if t[0] in "+-":
        t = t[1:]
    e10 = 0
    if "e" in t or "E" in t:
        t, ex = t.replace("E", "e").split("e")
        e10 = int(ex)
    if "." in t:
        ip, fp = t.split(".")
        e10 -= len(fp)
        t = ip + fp
    man = int(t) if t else 0
    if man == 0:
        return 0, EZ
    if e10 >= 0:
        v = man * 10 ** e10
        bl = v.bit_length()
        if bl <= NBb:
            return (-v if neg else v), 0
        sh = bl - NBb
        v = (v + (1 << (sh - 1))) >> sh
        return (-v if neg else v), sh
    den = 10 ** (-e10)
    sh = NBb + den.bit_length() - man.bit_length() + 1
    q = ((man << sh) + den // 2) // den
    ex = -sh
    while q.bit_length() > NBb:
        q = (q + 1) >> 1
        ex += 1
    return (-q if neg else q), ex


def from_strings(strs, L, shape=None):
    """decimal strings -> MPF (mpf.py:264-270)"""
    NBb = NBITS(L)
    ints, exps = [], []
    for t in strs:
        m_, e_ = _dec_to_bin(t, NBb)
        ints.append(m_)
        exps.append(e_)
    return from_ints(ints, exps, L, shape or (len(ints),))


def strings_to_host(strs, L):
    """decimal strings -> (limbs int32 (n, L) entry-major numpy, exps int64 numpy) (mpf.py:273-281)"""
    NBb = NBITS(L)
    nb = 4 * L
    n = len(strs)
    buf = bytearray(nb * n)
    exps = np.empty(n, dtype=np.int64)
    for i, t in enumerate(strs):
        m_, e_ = _dec_to_bin(t, NBb)
        buf[i * nb:(i + 1) * nb] = m_.to_bytes(nb, "little", signed=True)
        exps[i] = e_
    return np.frombuffer(bytes(buf), dtype=np.int32).reshape(n, L), exps


def host_to_mpf(limbs, exps, L, shape):
    """inverse of strings_to_host, on the GPU (normalised) (mpf.py:284-288)"""
    m = torch.from_numpy(np.ascontiguousarray(limbs.T)).to(dev)
    e = torch.from_numpy(exps).to(dev)
    return normalize(m, e, L).reshape(*shape)


def from_mp(vals, L, shape=None):
    """iterable of mpmath numbers / strings -> MPF (rounded to NB bits) (mpf.py:291-304)"""
    vals = list(vals)
    NBb = NBITS(L)
    ints, exps = [], []
    for v in vals:
        v = mp.mpf(v)
        if v == 0:
            ints.append(0)
            exps.append(EZ)
            continue
        sign, man, ex, bc = v._mpf_
        man = int(man)
        if bc > NBb:                                  # round to NB bits (half up on magnitude)
            sh = bc - NBb
            man = (man + (1 << (sh - 1))) >> sh
            ex += sh
        ints.append(-man if sign else man)
        exps.append(ex)
    return from_ints(ints, exps, L, shape or (len(vals),))


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
    return from_mp([v], L, (1,))


def eye(n, L, val=1):
    A = zeros((n, n), L)
    d = from_mp([val], L, (1,))
    idx = torch.arange(n, device=dev)
    A.m[:, idx, idx] = d.m[:, :1]
    A.e[idx, idx] = d.e[0]
    return A


# ------------------------------------------------------------------ reductions (host results)
def _top(A):
    """per-entry e + bitlen (log2 of magnitude, roughly); EZ for zeros (mpf.py:348-351)"""
    m, e = A.m.contiguous(), A.e.contiguous()
    out = torch.empty_like(e)
    call("sdpv_mpf_top", ptr(m), ptr(e), A.L, e.numel(), ptr(out), stream())
    return out


def total(A):
    """exact-ish sum of all entries -> mpmath mpf (error ~ 2^-NB * n * max) (mpf.py:354-366)"""
    L = A.L
    n = A.e.numel()
    if n == 0:
        return mp.mpf(0)
    t = int(A.e.max()) - 32
    if t < EZ // 2:
        return mp.mpf(0)
    s = (t - A.e.reshape(-1))
    W = shift_round(A.m.reshape(L, -1), s, L + 2)
    S = (W.to(torch.int64) & 0xFFFFFFFF).sum(1)
    top = W[-1].to(torch.int64).sum()
    S = S.cpu().tolist()
    top = int(top)
    v = sum(int(S[k]) << (32 * k) for k in range(L + 1)) + (top << (32 * (L + 1)))
    return mp.mpf((v, t)) if v else mp.mpf(0)


def dot(A, B):
    return total(mul(A, B))


def max_abs(A):
    """(mpf.py:372-381)"""
    if A.e.numel() == 0:
        return mp.mpf(0)
    tp = _top(A).reshape(-1)
    t = int(tp.max())
    if t <= EZ // 2:
        return mp.mpf(0)
    cand = torch.nonzero(tp >= t - 1).reshape(-1)
    Af = A.reshape(-1)
    vals = MPF(Af.m[:, cand], Af.e[cand]).shift(-t).to_float()
    i = int(cand[int(np.argmax(np.abs(vals)))])
    return abs(Af[i:i + 1].to_mp()[0])


def trace(A):
    n = A.shape[0]
    idx = torch.arange(n, device=dev)
    return total(MPF(A.m[:, idx, idx], A.e[idx, idx]))


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
        else:                                  # column panel: rows j0..j0+jw of op(B)^T
            Eb = EB.view(nb, N)[:, j0:j0 + jw].contiguous()
            off = j0 * bldr
            Bms = Bm.view(L, -1)[:, off:].contiguous()
            Bes = Be.view(-1)[off:].contiguous()
            R.split(Bms, Bes, Eb, L, Bes.numel(), nb, jw, bldb, bldr, bldk, k0, kc, kp, Rh, Rl)
        return Rh, Rl

    for j0, jw, acc in residue_gemm(R, fa, fb, nb, M, N, K, ACC_BUDGET):
        R.crt(acc, nb * M, jw, 0, jw, 1, outm=out_m, oute=out_e, L=L, plane=nb * M * N, ldo=N, jo=j0, Mb=M,
              erow=EA, ecol=EB, scaled=True)
    return MPF(out_m, out_e)


def symmetrize(A):
    """(A + A^T)/2"""
    return add(A, A.T).shift(-1)


# ------------------------------------------------------------------ Cholesky (equilibrated, batched)
def fx_exp(L):
    """exponent of the equilibrated fixed-point format: |x| <= 1 with 40 bits of integer headroom
    below the top limb (linalg.py:19)"""
    return -(32 * L - 40)


class Chol:
    """A = L L^T for one (n, n) or nb batched (nb, n, n) SPD matrices, with L = D^-1 Lt,
    D = diag(2^-k): Lt is fixed point at fx_exp(L), k integer per row (mpf.py:461-532).

    The equilibrated D A D (unit-order diagonal) is padded with identity to a multiple of bt and
    factored by the batched recursive Cholesky; solves go through the explicit tile inverses."""

    SOLVE_REFINE = int(os.environ.get("SDPB_GPU_SOLVE_REFINE", "0"))

    def __init__(self, A, bt=32, refine=1, dims=None):
        """dims: optional (nb,) true dimensions of batched blocks embedded in the top-left corner of
        A (zeros elsewhere); the rows beyond a block's dimension get the identity padding, as rows
        beyond n do, so each block factors exactly as it would alone."""
        from .batched import BMP, bchol
        self.two_d = A.e.dim() == 2
        if self.two_d:
            A = MPF(A.m[:, None], A.e[None])
        nb, n, _ = A.shape
        L = A.L
        self.nb, self.n, self.L, self.bt, self.refine = nb, n, L, bt, refine
        d = torch.arange(n, device=A.e.device)
        top_d = _top(A)[:, d, d]                                    # (nb, n)
        neg_d = A.m[-1][:, d, d] < 0
        live = None if dims is None else d[None, :] < torch.as_tensor(dims, device=A.e.device)[:, None]
        bad = neg_d | (top_d <= EZ // 2)
        if bool((bad if live is None else bad & live).any()):
            raise ValueError("matrix is not positive definite (diagonal)")
        k = torch.div(top_d + 1, 2, rounding_mode="floor")          # A_ii < 2^(2 k_i)
        if live is not None:
            k = torch.where(live, k, torch.zeros_like(k))
        self.k = k
        self.npad = npad = max(bt, -(-n // bt) * bt)
        efx = fx_exp(L)
        # (D A D)_ij = m_ij 2^(e_ij - k_i - k_j), re-rounded to the fixed-point exponent
        s = efx - (A.e - k[:, :, None] - k[:, None, :])
        s = torch.where(A.e == EZ, torch.full_like(s, 1 << 30), s).reshape(-1)
        Af = shift_round(A.m.reshape(L, -1), s, L).view(L, nb, n, n)
        P = torch.zeros((L, nb, npad, npad), dtype=torch.int32, device=A.m.device)
        P[:, :, :n, :n] = Af
        one = 1 << (-efx)                                           # identity padding: 1 = 2^-efx units
        words = [(one >> (32 * w)) & 0xFFFFFFFF for w in range(L)]
        words = [v - (1 << 32) if v >> 31 else v for v in words]
        if npad > n:
            pad = torch.arange(n, npad, device=A.m.device)
            for w in range(L):
                P[w][:, pad, pad] = words[w]
        if live is not None and not bool(live.all()):
            bi, ri = (~live).nonzero(as_tuple=True)
            for w in range(L):
                P[w][bi, ri, ri] = words[w]
        self.Lt, self.tiles = bchol(BMP(P, efx), bt, refine)

    def _to_bmp(self, B):
        """right-hand sides (nb, n, m) -> fixed point with one exponent per (block, column)"""
        from .batched import BMP
        L = self.L
        Bm, cE = _align_rows(B.T)                                   # (L, nb, m, n), (nb, m)
        P = torch.zeros((L, self.nb, self.npad, B.shape[-1]), dtype=torch.int32, device=Bm.device)
        P[:, :, :self.n] = Bm.transpose(2, 3)
        return BMP(P, cE)

    def _from_bmp(self, P):
        L, m = self.L, P.shape[2]
        X = P.limbs[:, :, :self.n].reshape(L, -1).contiguous()
        ein = P.expt()[:, None, :].expand(self.nb, self.n, m).reshape(-1).contiguous()
        return normalize(X, ein, L).reshape(self.nb, self.n, m)

    def _io(self, B, f):
        if self.two_d:
            r = f(MPF(B.m[:, None], B.e[None]))
            return MPF(r.m[:, 0], r.e[0])
        return f(B)

    def solve_lower(self, B):
        """L^-1 B = Lt^-1 (D B)"""
        from .batched import btrsm_left

        def f(B):
            P = btrsm_left(self.Lt, self._to_bmp(B.shift(-self.k[:, :, None])), self.tiles, self.bt,
                           self.SOLVE_REFINE)
            return self._from_bmp(P)
        return self._io(B, f)

    def solve_upper(self, B):
        """L^-T B = D (Lt^-T B)"""
        from .batched import btrsm_left_T

        def f(B):
            P = btrsm_left_T(self.Lt, self._to_bmp(B), self.tiles, self.bt, self.SOLVE_REFINE)
            return self._from_bmp(P).shift(-self.k[:, :, None])
        return self._io(B, f)

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
        if self.acc is None:
            return
        q = self._crt()
        self.part = q if self.part is None else add(self.part, q)
        self.acc = None
        self.flushes += 1

    def next_reference(self, margin=4):
        """reference exponents for the next accumulation (columns never seen: none)"""
        return torch.where(self.Tmax > EZ // 2, self.Tmax + margin, self.Tmax)

    def precision_loss(self):
        """max over columns of (reference - actual column maximum), in bits"""
        ok = (self.Tmax > EZ // 2) & (self.E > EZ // 2)
        return int((self.E - self.Tmax)[ok].max()) if bool(ok.any()) else 0

    def result(self):
        if self.acc is None:
            return self.part
        q = self._crt()
        self.acc = None
        return q if self.part is None else add(self.part, q)
