"""Multiprecision matrices in limb-plane layout (mirror of sdpvibe/mpmat.py).

An MP matrix holds limbs: int32 tensor of shape (L, rows, cols) -- limb plane k contains the k-th
32-bit limb (little endian, two's complement) of every entry -- and exponents:
    value[i, j] = int(limbs[:, i, j]) * 2**exp[j]       (exp: (cols,) int64, or a Python int e0)
"""
import torch

from .limbops import add_same, bitlen_words, shift_round
from .rns import KMAX_LAUNCH, RNS, dev

# device memory budgets of one gemm_nt call (bytes); 288 GB of HBM leaves room for large panels
ACC_BUDGET = 4 << 30
RES_BUDGET = 4 << 30


class MP:
    """e0: Python int when the exponent is uniform (avoids device syncs), else None (mpmat.py:14-56)"""
    __slots__ = ("limbs", "_exp", "e0")

    def __init__(self, limbs, exp):
        self.limbs = limbs
        if isinstance(exp, int):
            self.e0, self._exp = exp, None
        elif torch.is_tensor(exp) and exp.dim() == 0:
            self.e0, self._exp = None, exp.expand(limbs.shape[2]).clone()
        else:
            self.e0, self._exp = None, exp

    @property
    def exp(self):
        if self._exp is None:
            self._exp = torch.full((self.limbs.shape[2],), self.e0, dtype=torch.int64, device=self.limbs.device)
        return self._exp

    @property
    def shape(self):
        return tuple(self.limbs.shape[1:])

    @property
    def L(self):
        return self.limbs.shape[0]

    def t(self):
        """transpose; requires a single exponent"""
        return MP(self.limbs.transpose(1, 2).contiguous(), self.uniform_exp())

    def uniform_exp(self):
        if self.e0 is not None:
            return self.e0
        e = int(self._exp[0])
        assert bool((self._exp == e).all()), "matrix has per-column exponents"
        self.e0 = e
        return e

    def __getitem__(self, idx):
        r, c = idx
        if self.e0 is not None:
            return MP(self.limbs[:, r, c].contiguous(), self.e0)
        return MP(self.limbs[:, r, c].contiguous(), self._exp[c].contiguous())

    def flat(self):
        """(L, rows*cols) view"""
        return self.limbs.reshape(self.L, -1)


def transpose_rows(B):
    """B (n x m) with per-column exponents -> limbs of B^T (m x n) and the row exponents (m,)"""
    return MP(B.limbs.transpose(1, 2).contiguous(), 0), B.exp


def headroom_bits(L):
    return 32 * L - 32          # target mantissa size for dynamic rounding


# ------------------------------------------------------------------ host conversions (tests)
def from_ints(rows, exp, L):
    """rows: list of lists of Python ints (mantissas) -> MP (mpmat.py:68-80)"""
    R, C = len(rows), len(rows[0])
    mask = (1 << 32) - 1
    flat = []
    for row in rows:
        for v in row:
            assert abs(v) < 1 << (32 * L - 1)
            u = v % (1 << (32 * L))
            flat.append([((u >> (32 * k)) & mask) for k in range(L)])
    arr = torch.tensor(flat, dtype=torch.int64).view(R, C, L).permute(2, 0, 1).contiguous()
    arr = torch.where(arr >= 1 << 31, arr - (1 << 32), arr).to(torch.int32)
    return MP(arr.to(dev), exp)


def to_ints(A):
    """MP -> list of lists of Python int mantissas, list of column exponents (mpmat.py:83-98)"""
    L = A.L
    a = (A.limbs.cpu().to(torch.int64) % (1 << 32)).permute(1, 2, 0)
    R, C = A.shape
    out = []
    for i in range(R):
        row = []
        for j in range(C):
            v = 0
            for k in reversed(range(L)):
                v = (v << 32) | int(a[i, j, k])
            if v >= 1 << (32 * L - 1):
                v -= 1 << (32 * L)
            row.append(v)
        out.append(row)
    return out, A.exp.cpu().tolist()


# ------------------------------------------------------------------ re-rounding
def _exp_vec(e, C, device=None):
    if torch.is_tensor(e):
        return e if e.dim() == 1 else e.expand(C)
    return torch.full((C,), int(e), dtype=torch.int64, device=device or dev)


def rescale(A, new_exp, Lout=None):
    """re-round A to exponent(s) new_exp (int, 0-d or (cols,) tensor), optionally to Lout limbs
    (mpmat.py:108-117)"""
    R, C = A.shape
    Lo = Lout or A.L
    if isinstance(new_exp, int) and A.e0 is not None:
        s = torch.full((R * C,), new_exp - A.e0, dtype=torch.int64, device=A.limbs.device)
        return MP(shift_round(A.flat(), s, Lo).view(Lo, R, C), new_exp)
    ne = _exp_vec(new_exp, C, A.limbs.device)
    s = (ne - A.exp)[None, :].expand(R, C).reshape(-1)
    return MP(shift_round(A.flat(), s, Lo).view(Lo, R, C), ne.clone())


def bitlen_limbs(X):
    return bitlen_words(X.reshape(X.shape[0], -1))


def normalize(A, mode="mat"):
    R, C = A.shape
    top = bitlen_limbs(A.limbs).view(R, C) + A.exp[None, :]
    t = top.max() if mode == "mat" else top.amax(0)
    return rescale(A, t - headroom_bits(A.L))


def neg(A):
    R, C = A.shape
    return MP(add_same(torch.zeros_like(A.flat()), A.flat(), -1).view(A.L, R, C),
              A.e0 if A.e0 is not None else A.exp.clone())


def add(A, B, alpha=1, out_exp=None):
    """A + alpha*B (alpha = +-1). Exact when exponents agree and out_exp equals them.
    out_exp='align': floating-point style (round both to the larger exponent per column)
    (mpmat.py:135-156)."""
    assert A.shape == B.shape and A.L == B.L
    R, C = A.shape
    L = A.L
    if isinstance(out_exp, int) and A.e0 == out_exp and B.e0 == out_exp:
        return MP(add_same(A.flat(), B.flat(), alpha).view(L, R, C), out_exp)
    if out_exp == "align":
        e = torch.maximum(A.exp, B.exp) + 1
        a = rescale(A, e)
        b = rescale(B, e)
        return MP(add_same(a.flat(), b.flat(), alpha).view(L, R, C), e)
    e = torch.minimum(A.exp, B.exp)
    d = int(max((A.exp - e).max(), (B.exp - e).max()))
    Lw = L + (d + 31) // 32 + 2
    a = rescale(A, e, Lw)
    b = rescale(B, e, Lw)
    W = MP(add_same(a.flat(), b.flat(), alpha).view(Lw, R, C), e.clone())
    if out_exp is None or out_exp in ("mat", "col"):
        top = bitlen_limbs(W.limbs).view(R, C) + W.exp[None, :]
        top = top.amax(0) if out_exp == "col" else top.max()
        out_exp = top - headroom_bits(L)
    return rescale(W, out_exp, L)


# ------------------------------------------------------------------ GEMM
_ctx = {}


def ctx(L):
    """RNS context of L-limb operands on the current device (mpmat.py:163-166)"""
    key = (L, torch.cuda.current_device())
    if key not in _ctx:
        _ctx[key] = RNS(L)
    return _ctx[key]


def residue_gemm(R, A_desc, B_desc, nb, M, N, K, acc_rows_budget=ACC_BUDGET):
    """Generator of (j0, jw, acc) panels: acc (n, nb*M, jw) int16 holding the CRT coefficients
    (centred residue * (M/m)^-1 mod m, for crt(..., scaled=True)) of the exact products over the
    whole inner dimension K. A_desc/B_desc: callables (k0, kc, kp, j0, jw) ->
    (Rh, Rl) residues (n, nb*rows, kp) of op(A) rows / op(B)^T rows j0..j0+jw."""
    n = R.n
    w = N if n * nb * M * N * 2 <= acc_rows_budget else max(64, acc_rows_budget // (n * nb * M * 2) // 64 * 64)
    assert K <= R.kmax, f"inner dimension {K} exceeds the RNS bound {R.kmax}"
    for j0 in range(0, N, w):
        jw = min(w, N - j0)
        acc = torch.empty((n, nb * M, jw), dtype=torch.int16, device=dev)
        per_k = 2 * n * nb * (M + jw) * 2
        kc = max(64, min(KMAX_LAUNCH, RES_BUDGET // max(per_k, 1) // 64 * 64))
        first = True
        if K == 0:
            acc.zero_()
        for k0 in range(0, K, kc):
            kk = min(kc, K - k0)
            kp = R.kp(kk)
            Ah, Al = A_desc(k0, kk, kp, j0, jw)
            Bh, Bl = B_desc(k0, kk, kp, j0, jw)
            R.product(Ah, Al, Bh, Bl, acc, nb, M, jw, nb * M * jw, accumulate=not first, scaled=True)
            first = False
            del Ah, Al, Bh, Bl
        yield j0, jw, acc


def _limb_desc(R, X, rows, K, nb=1):
    """residue callable for a fixed-point (L, nb*rows, K) limb tensor (no shift)"""
    L = X.shape[0]

    def f(k0, kc, kp, j0, jw, X=X):
        r0, rn = (0, rows) if j0 is None else (j0, jw)
        Xs = X.view(L, nb, rows, K)
        base = Xs[:, :, r0:r0 + rn] if (r0, rn) != (0, rows) else Xs
        base = base.contiguous()
        Rh = torch.empty((R.n, kp // 16, nb * rn, 16), dtype=torch.int8, device=dev)
        Rl = torch.empty_like(Rh)
        R.split(base, None, None, L, nb * rn * K, nb, rn, rn * K, K, 1, k0, kc, kp, Rh, Rl)
        return Rh, Rl
    return f


def gemm_nt(A, Bt, out_exp="col", kc=4096, colblock=None, bt_rowexp=None):
    """C = A @ Bt^T, exact, then rounded.  A: (M, K) single exponent; Bt: (N, K) single exponent,
    or per-row exponents bt_rowexp (N,).  out_exp: 'col' (dynamic per column), 'mat' (dynamic
    common), or an int / (N,) tensor of target exponents (mpmat.py:169-202)."""
    M, K = A.shape
    N, K2 = Bt.shape
    assert K == K2
    L = A.L
    R = ctx(L)
    eA = A.uniform_exp()
    e_in = (eA + bt_rowexp) if bt_rowexp is not None else eA + Bt.uniform_exp()
    same = A.limbs.data_ptr() == Bt.limbs.data_ptr()
    fa = _limb_desc(R, A.limbs.contiguous(), M, K)
    fb = _limb_desc(R, Bt.limbs.contiguous(), N, K)
    cache = {}

    def fa0(k0, kc_, kp, j0, jw):
        cache["a"] = fa(k0, kc_, kp, None, None)
        return cache["a"]

    def fb0(k0, kc_, kp, j0, jw):
        if same and (j0, jw) == (0, N):
            return cache["a"]
        return fb(k0, kc_, kp, j0, jw)

    out = torch.empty((L, M, N), dtype=torch.int32, device=dev)
    fixed = isinstance(e_in, int) and isinstance(out_exp, int)
    exps = None if fixed else torch.empty((N,), dtype=torch.int64, device=dev)
    if not fixed:
        e_in = _exp_vec(e_in, N)
    for j0, jw, acc in residue_gemm(R, fa0, fb0, 1, M, N, K):
        _crt_round_panel(R, acc, M, jw, j0, N, e_in, out_exp, L, out, exps, fixed)
    if fixed:
        return MP(out, out_exp)
    C = MP(out, exps)
    if out_exp == "mat":                                           # common exponent (one extra rounding)
        C = rescale(C, exps.max())
    return C


def _crt_round_panel(R, acc, M, jw, j0, N, e_in, out_exp, L, out, exps, fixed):
    """_crt_round of one column panel (mpmat.py:205-236): exact integers -> L limbs at the output
    exponents."""
    if fixed:
        s = torch.full((N,), out_exp - e_in, dtype=torch.int64, device=dev)
        R.crt(acc, M, jw, 0, jw, 2, outm=out, L=L, plane=M * N, ldo=N, jo=j0, Mb=M, scol=s, scaled=True,
              flag=_flag())
        return
    Wd, bl, Bq = R.crt_block(acc, jw, 0, jw, M, scaled=True)
    if isinstance(out_exp, str):
        sj = bl.view(M, jw).amax(0) - headroom_bits(L)
        ej = e_in[j0:j0 + jw] + sj
    else:
        ej = _exp_vec(out_exp, N)[j0:j0 + jw]
        sj = ej - e_in[j0:j0 + jw]
    s = sj[None, :].expand(M, jw).reshape(-1)
    exps[j0:j0 + jw] = ej
    out[:, :, j0:j0 + jw] = shift_round(Wd, s, L).view(L, M, jw)


def _flag():
    from .limbops import _errbuf
    return _errbuf()
