"""Multiprecision matrices in limb-plane layout (mirror of sdpvibe/mpmat.py).

An MP matrix holds limbs: int32 tensor of shape (L, rows, cols) -- limb plane k contains the k-th
32-bit limb (little endian, two's complement) of every entry -- and exponents:
    value[i, j] = int(limbs[:, i, j]) * 2**exp[j]       (exp: (cols,) int64, or a Python int e0)
"""
import numpy as np
import torch

from .limbops import add_same, bitlen_words, shift_round
from .rns import KMAX_LAUNCH, RNS, dev

# device memory budgets of one gemm_nt call (bytes); 288 GB of HBM leaves room for large panels
ACC_BUDGET = 4 << 30
RES_BUDGET = 4 << 30


class MP:
    """Fixed-point or block-floating matrix: limbs (L, R, C) and either one exponent for the whole
    matrix (a Python int: no device round trip to read it) or one per column ((C,) int64 tensor).
    `e0` is the uniform exponent when it is known on the host, else None."""
    __slots__ = ("limbs", "e0", "_cols")

    def __init__(self, limbs, exp):
        self.limbs = limbs
        self.e0, self._cols = None, None
        if isinstance(exp, int):
            self.e0 = exp
        elif exp.dim() == 0:                           # one exponent held on the device
            self._cols = exp.reshape(1).repeat(limbs.shape[2])
        else:
            self._cols = exp

    @property
    def L(self):
        return self.limbs.shape[0]

    @property
    def shape(self):
        return tuple(self.limbs.shape[1:])

    @property
    def exp(self):
        """per-column exponents as a (C,) tensor (materialised from e0 when uniform)"""
        if self._cols is None:
            self._cols = torch.full((self.limbs.shape[2],), self.e0, dtype=torch.int64, device=self.limbs.device)
        return self._cols

    def uniform_exp(self):
        """the common exponent as an int (reads the device once when only the columns are known)"""
        if self.e0 is None:
            lo, hi = (int(v) for v in torch.aminmax(self._cols))
            if lo != hi:
                raise ValueError("matrix has per-column exponents")
            self.e0 = lo
        return self.e0

    def t(self):
        """transpose (needs a common exponent)"""
        return MP(self.limbs.transpose(1, 2).contiguous(), self.uniform_exp())

    def __getitem__(self, idx):
        rows, cols = idx
        part = self.limbs[:, rows, cols].contiguous()
        return MP(part, self.e0 if self.e0 is not None else self._cols[cols].contiguous())

    def flat(self):
        return self.limbs.reshape(self.L, -1)


def transpose_rows(B):
    """(limbs of B^T as an exponent-0 MP, B's column exponents): B^T as the row-scaled operand of
    gemm_nt"""
    return MP(B.limbs.transpose(1, 2).contiguous(), 0), B.exp


def headroom_bits(L):
    """mantissa size that dynamic roundings aim at: 32 bits below the top of L limbs"""
    return 32 * L - 32


# ------------------------------------------------------------------ host conversions (tests)
def from_ints(rows, exp, L):
    """list of rows of Python ints (|v| < 2^(32L-1)) -> MP with exponent(s) exp"""
    R, C = len(rows), len(rows[0])
    vals = [v for row in rows for v in row]
    if any(abs(v) >= 1 << (32 * L - 1) for v in vals):
        raise OverflowError("value does not fit in L limbs")
    raw = b"".join(v.to_bytes(4 * L, "little", signed=True) for v in vals)
    words = np.frombuffer(raw, dtype=np.int32).reshape(R, C, L).transpose(2, 0, 1)
    return MP(torch.from_numpy(np.array(words)).to(dev), exp)


def to_ints(A):
    """MP -> (rows of Python int mantissas, list of column exponents)"""
    R, C = A.shape
    words = A.limbs.permute(1, 2, 0).contiguous().cpu().numpy()
    rows = [[int.from_bytes(words[i, j].tobytes(), "little", signed=True) for j in range(C)] for i in range(R)]
    return rows, A.exp.cpu().tolist()


# ------------------------------------------------------------------ re-rounding and additions
def _col_exps(e, C, device):
    """exponent argument (int, 0-d or (C,) tensor) as a (C,) tensor"""
    if not torch.is_tensor(e):
        return torch.full((C,), int(e), dtype=torch.int64, device=device)
    return e.reshape(-1).expand(C) if e.numel() == 1 else e


def rescale(A, new_exp, Lout=None):
    """A re-rounded (shift_round: half toward +inf) to the exponent(s) new_exp, into Lout limbs"""
    R, C = A.shape
    Lo = Lout or A.L
    dev_ = A.limbs.device
    if isinstance(new_exp, int) and A.e0 is not None:
        shift = torch.full((R * C,), new_exp - A.e0, dtype=torch.int64, device=dev_)
        return MP(shift_round(A.flat(), shift, Lo).view(Lo, R, C), new_exp)
    target = _col_exps(new_exp, C, dev_)
    shift = (target - A.exp).repeat(R)
    return MP(shift_round(A.flat(), shift, Lo).view(Lo, R, C), target.clone())


def bitlen_limbs(X):
    return bitlen_words(X.reshape(X.shape[0], -1))


def _tops(A):
    """(R, C) binary exponent of the top bit of every entry"""
    R, C = A.shape
    return bitlen_limbs(A.limbs).view(R, C) + A.exp[None, :]


def normalize(A, mode="mat"):
    """re-round so the largest entry (mode 'mat') or each column's largest (mode 'col') has
    headroom_bits(L) bits"""
    t = _tops(A)
    top = t.max() if mode == "mat" else t.amax(0)
    return rescale(A, top - headroom_bits(A.L))


def neg(A):
    """-A, exact (two's complement wrap)"""
    R, C = A.shape
    zero = torch.zeros_like(A.flat())
    return MP(add_same(zero, A.flat(), -1).view(A.L, R, C), A.e0 if A.e0 is not None else A.exp.clone())


def add(A, B, alpha=1, out_exp=None):
    """A + alpha B (alpha = +-1).
    out_exp int equal to both operands' common exponent: exact limb addition.
    out_exp 'align': both operands rounded to one more than the larger exponent of each column, then
    added (the floating-point style sum).
    otherwise: exact sum at the smaller exponent of each column in enough limbs, then one rounding to
    out_exp (int or tensor), or to the dynamic exponent of the matrix (None, 'mat') or of each
    column ('col')."""
    if A.shape != B.shape or A.L != B.L:
        raise ValueError("operands differ in shape or limb count")
    R, C = A.shape
    L = A.L
    if isinstance(out_exp, int) and A.e0 == out_exp and B.e0 == out_exp:
        return MP(add_same(A.flat(), B.flat(), alpha).view(L, R, C), out_exp)
    if out_exp == "align":
        e = torch.maximum(A.exp, B.exp) + 1
        s = add_same(rescale(A, e).flat(), rescale(B, e).flat(), alpha)
        return MP(s.view(L, R, C), e)
    low = torch.minimum(A.exp, B.exp)
    spread = int(torch.maximum(A.exp - low, B.exp - low).max())
    Lw = L + 2 + (spread + 31) // 32                     # room for the larger operand and the carry
    wide = MP(add_same(rescale(A, low, Lw).flat(), rescale(B, low, Lw).flat(), alpha).view(Lw, R, C), low.clone())
    if out_exp is None or isinstance(out_exp, str):
        t = _tops(wide)
        out_exp = (t.amax(0) if out_exp == "col" else t.max()) - headroom_bits(L)
    return rescale(wide, out_exp, L)


# ------------------------------------------------------------------ GEMM
_ctx = {}


def ctx(L):
    """RNS context of L-limb operands on the current device (mpmat.py:163-166)"""
    key = (L, torch.cuda.current_device())
    if key not in _ctx:
        _ctx[key] = RNS(L)
    return _ctx[key]


def residue_gemm(R, A_desc, B_desc, nb, M, N, K, acc_rows_budget=None):
    """Generator of (j0, jw, acc) panels: acc (n, nb*M, jw) int16 holding the CRT coefficients
    (centred residue * (M/m)^-1 mod m, for crt(..., scaled=True)) of the exact products over the
    whole inner dimension K. A_desc/B_desc: callables (k0, kc, kp, j0, jw) ->
    (Rh, Rl) residues (n, nb*rows, kp) of op(A) rows / op(B)^T rows j0..j0+jw."""
    n = R.n
    acc_rows_budget = ACC_BUDGET if acc_rows_budget is None else acc_rows_budget   # read per call
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
        Rh = torch.empty((R.n, kp // 16, nb * rn, 16), dtype=torch.int8, device=dev)
        Rl = torch.empty_like(Rh)
        # a panel is read in place: rows r0.. of every block, offset r0*K into each limb plane
        R.split(X, None, None, L, nb * rows * K, nb, rn, rows * K, K, 1, k0, kc, kp, Rh, Rl, off=r0 * K)
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
        e_in = _col_exps(e_in, N, dev)
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
        ej = _col_exps(out_exp, N, dev)[j0:j0 + jw]
        sj = ej - e_in[j0:j0 + jw]
    s = sj[None, :].expand(M, jw).reshape(-1)
    exps[j0:j0 + jw] = ej
    out[:, :, j0:j0 + jw] = shift_round(Wd, s, L).view(L, M, jw)


def _flag():
    from .limbops import _errbuf
    return _errbuf()
