"""Batched multiprecision linear algebra: many independent equal-sized blocks per launch (mirror of
sdpvibe/batched.py).

A BMP holds a stack of nb matrices:
    limbs: int32 (L, nb, R, C) limb planes;  exp: Python int (uniform) or (nb, C) int64 tensor.
"""
import torch

from ._lib import call, ptr, stream
from .limbops import _errbuf, add_same, shift_round
from .mpmat import ACC_BUDGET, ctx, headroom_bits, residue_gemm
from .rns import dev


class BMP:
    """(batched.py:16-42)"""
    __slots__ = ("limbs", "exp")

    def __init__(self, limbs, exp):
        self.limbs, self.exp = limbs, exp            # exp: int or (nb, C) tensor

    @property
    def shape(self):
        return tuple(self.limbs.shape[1:])

    @property
    def L(self):
        return self.limbs.shape[0]

    def expt(self):
        nb, R, C = self.shape
        if isinstance(self.exp, int):
            return torch.full((nb, C), self.exp, dtype=torch.int64, device=self.limbs.device)
        return self.exp

    def __getitem__(self, idx):
        r, c = idx
        e = self.exp if isinstance(self.exp, int) else self.exp[:, c].contiguous()
        return BMP(self.limbs[:, :, r, c].contiguous(), e)

    def t(self):
        assert isinstance(self.exp, int)
        return BMP(self.limbs.transpose(2, 3).contiguous(), self.exp)

    def flat(self):
        return self.limbs.reshape(self.L, -1)


def brescale(A, new_exp, Lout=None):
    """re-round every block to exponent(s) new_exp (int or (nb, C)) (batched.py:45-52)"""
    nb, R, C = A.shape
    Lo = Lout or A.L
    if isinstance(new_exp, int) and isinstance(A.exp, int):
        s = torch.full((nb * R * C,), new_exp - A.exp, dtype=torch.int64, device=A.limbs.device)
        return BMP(shift_round(A.flat(), s, Lo).view(Lo, nb, R, C), new_exp)
    ne = new_exp if torch.is_tensor(new_exp) else torch.full((nb, C), int(new_exp), dtype=torch.int64,
                                                              device=A.limbs.device)
    s = (ne - A.expt())[:, None, :].expand(nb, R, C).reshape(-1)
    return BMP(shift_round(A.flat(), s, Lo).view(Lo, nb, R, C), ne.clone())


def badd(A, B, alpha=1, out_exp=None):
    """A + alpha*B: exact when both exponents equal out_exp (int), else 'align' (both rounded to
    max(ea, eb) + 1 per column, then added) (batched.py:82-92)"""
    nb, R, C = A.shape
    L = A.L
    if isinstance(out_exp, int) and A.exp == out_exp and B.exp == out_exp:
        return BMP(add_same(A.flat(), B.flat(), alpha).view(L, nb, R, C), out_exp)
    assert out_exp == "align"
    out = torch.empty((L, nb, R, C), dtype=torch.int32, device=A.limbs.device)
    e = torch.empty((nb, C), dtype=torch.int64, device=A.limbs.device)
    Al, Bl = A.flat().contiguous(), B.flat().contiguous()
    Ae, Be = A.expt().contiguous(), B.expt().contiguous()
    call("sdpv_badd_align", ptr(Al), ptr(Bl), ptr(Ae), ptr(Be), ptr(out), ptr(e), L, nb, R, C, int(alpha), stream())
    return BMP(out, e)


def _residue_desc(R, X, nb, rows, K):
    """residues of a fixed-point (L, nb, rows, K) limb tensor, all primes; panels of rows j0..j0+jw"""
    L = X.shape[0]

    def f(k0, kc, kp, j0, jw, X=X):
        if j0 is None or (j0, jw) == (0, rows):
            base, rn = X, rows
        else:
            base, rn = X[:, :, j0:j0 + jw].contiguous(), jw
        Rh = torch.empty((R.n, kp // 16, nb * rn, 16), dtype=torch.int8, device=X.device)
        Rl = torch.empty_like(Rh)
        R.split(base, None, None, L, nb * rn * K, nb, rn, rn * K, K, 1, k0, kc, kp, Rh, Rl)
        return Rh, Rl
    return f


def bgemm_nt(A, Bt, out_exp="col", kc=4096, bt_rowexp=None):
    """C[b] = A[b] @ Bt[b]^T for all b (exact, then rounded). A: uniform exponent per block (int or
    (nb, C) tensor read at column 0). Bt: int exponent or per-row exponents bt_rowexp (nb, N).
    out_exp: 'col' (dynamic per (block, column)), 'mat' (dynamic per block), or int (fixed)
    (batched.py:165-197)."""
    nb, M, K = A.shape
    nb2, N, K2 = Bt.shape
    assert nb == nb2 and K == K2
    L = A.L
    R = ctx(L)
    eA = A.exp if isinstance(A.exp, int) else A.exp[:, 0]
    Al = A.limbs.contiguous()
    Bl = Bt.limbs.contiguous()
    same = Al.data_ptr() == Bl.data_ptr()
    fa = _residue_desc(R, Al, nb, M, K)
    fb = _residue_desc(R, Bl, nb, N, K)
    cache = {}

    def fa0(k0, kc_, kp, j0, jw):
        cache["a"] = fa(k0, kc_, kp, None, None)
        return cache["a"]

    def fb0(k0, kc_, kp, j0, jw):
        if same and (j0, jw) == (0, N):
            return cache["a"]
        return fb(k0, kc_, kp, j0, jw)

    eB = Bt.exp if bt_rowexp is None else bt_rowexp
    if isinstance(eA, int) and isinstance(eB, int):
        e_in = eA + eB
    else:
        eA_t = eA if torch.is_tensor(eA) else torch.full((nb,), eA, dtype=torch.int64, device=dev)
        eB_t = eB if torch.is_tensor(eB) else torch.full((nb, N), eB, dtype=torch.int64, device=dev)
        if eB_t.dim() == 1:
            eB_t = eB_t[:, None].expand(nb, N)
        e_in = (eA_t[:, None] + eB_t).contiguous()
    out = torch.empty((L, nb, M, N), dtype=torch.int32, device=dev)
    fixed = isinstance(e_in, int) and isinstance(out_exp, int)
    exps = None if fixed else torch.empty((nb, N), dtype=torch.int64, device=dev)
    for j0, jw, acc in residue_gemm(R, fa0, fb0, nb, M, N, K, ACC_BUDGET):
        _bcrt_round_panel(R, acc, nb, M, N, j0, jw, e_in, out_exp, L, out, exps)
    if fixed or isinstance(out_exp, int):
        return BMP(out, out_exp)
    C = BMP(out, exps)
    if out_exp == "mat":
        C = brescale(C, exps.amax(1, keepdim=True).expand(nb, N).contiguous())
    return C


def _bcrt_round_panel(R, acc, nb, M, N, j0, jw, e_in, out_exp, L, out, exps):
    """_bcrt_round (batched.py:200-231) of one column panel: exact products -> L limbs"""
    rows = nb * M
    if isinstance(out_exp, int):
        if isinstance(e_in, int):
            s = torch.full((nb * N,), out_exp - e_in, dtype=torch.int64, device=dev)
        else:
            s = (out_exp - e_in).reshape(-1).contiguous()
            exps[:, j0:j0 + jw] = out_exp
        R.crt(acc, rows, jw, 0, jw, 2, outm=out, L=L, plane=nb * M * N, ldo=N, jo=j0, Mb=M, scol=s, flag=_errbuf(),
              scaled=True)
        return
    Wd, bl, Bq = R.crt_block(acc, jw, 0, jw, rows, True)               # entries (b*M + i)*jw + jj
    ein = e_in if torch.is_tensor(e_in) else torch.full((nb, N), e_in, dtype=torch.int64, device=dev)
    sj = bl.view(nb, M, jw).amax(1) - headroom_bits(L)             # (nb, jw)
    ej = ein[:, j0:j0 + jw] + sj
    s = sj[:, None, :].expand(nb, M, jw).reshape(-1)
    exps[:, j0:j0 + jw] = ej
    out[:, :, :, j0:j0 + jw] = shift_round(Wd, s, L).view(L, nb, M, jw)


# ------------------------------------------------------------------ batched Cholesky / TRSM
class BTile:
    """factor of a diagonal tile: Lmp (the factor, fixed point) and X (explicit inverse, one
    exponent per block)"""
    __slots__ = ("X", "Lmp")


def bchol_tile(A, LI=24, threads=256):
    """A: BMP (nb, b, b) fixed point at exponent -(32L-40) -> (L BMP, BTile with the inverse X)
    (batched.py:243-262)"""
    from .basecase import tile_factor
    from .limbops import bitlen_words
    nb, b, _ = A.shape
    L = A.L
    e = A.exp
    WL, F = L + LI, 32 * L
    assert isinstance(e, int) and e == -(32 * L - 40), "tiles enter at the fixed-point exponent"
    Lout_l, Xs, _ = tile_factor(A.limbs, b)
    Lout = BMP(Lout_l, e)
    top = bitlen_words(Xs.reshape(WL, -1)).view(nb, b * b).amax(1) - F
    ex = (top - headroom_bits(L))[:, None].expand(nb, b).contiguous()
    t = BTile()
    t.Lmp = Lout
    t.X = brescale(BMP(Xs, torch.full((nb, b), -F, dtype=torch.int64, device=Xs.device)), ex, L)
    return Lout, t


def _split(n, bt):
    return (n // 2 + bt - 1) // bt * bt


def _ntiles(n, bt):
    if n <= bt:
        return 1
    n1 = _split(n, bt)
    return _ntiles(n1, bt) + _ntiles(n - n1, bt)


def bchol(A, bt=32, refine=1):
    """batched Cholesky of nb SPD equilibrated fixed-point blocks (nb, n, n) -> (L, tiles)
    (batched.py:274-288)"""
    nb, n, _ = A.shape
    L = A.L
    e = A.exp
    if n <= bt:
        assert n == bt
        Lt, t = bchol_tile(A)
        return Lt, [t]
    n1 = _split(n, bt)
    L11, t1 = bchol(A[:n1, :n1], bt, refine)
    L21 = btrsm_right(A[n1:, :n1], L11, t1, bt, refine)
    S = badd(A[n1:, n1:], bgemm_nt(L21, L21, out_exp=e), -1, out_exp=e)
    L22, t2 = bchol(S, bt, refine)
    out = torch.zeros((L, nb, n, n), dtype=torch.int32, device=A.limbs.device)
    out[:, :, :n1, :n1] = L11.limbs
    out[:, :, n1:, :n1] = L21.limbs
    out[:, :, n1:, n1:] = L22.limbs
    return BMP(out, e), t1 + t2


def btrsm_right(B, Lf, tiles, bt=32, refine=1):
    """X with X L^T = B, batched; fixed point (batched.py:291-305)"""
    nb, m, n = B.shape
    e = Lf.exp
    if n <= bt:
        t = tiles[0]
        X = bgemm_nt(B, t.X, out_exp=e, bt_rowexp=t.X.exp)
        for _ in range(refine):
            Rr = badd(B, bgemm_nt(X, t.Lmp, out_exp=e), -1, out_exp=e)
            X = badd(X, bgemm_nt(Rr, t.X, out_exp=e, bt_rowexp=t.X.exp), 1, out_exp=e)
        return X
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    X1 = btrsm_right(B[:, :n1], Lf[:n1, :n1], tiles[:k1], bt, refine)
    B2 = badd(B[:, n1:], bgemm_nt(X1, Lf[n1:, :n1], out_exp=e), -1, out_exp=e)
    X2 = btrsm_right(B2, Lf[n1:, n1:], tiles[k1:], bt, refine)
    return BMP(torch.cat([X1.limbs, X2.limbs], 3), e)


def _tr(B):
    """limbs of B^T per block, for use as the 'Bt' operand (row exponents passed separately)"""
    return BMP(B.limbs.transpose(2, 3).contiguous(), 0)


def _bleaf(X, Lm, B, refine):
    """P = Lm^{-1} B with the explicit tile inverse X (+ refinement steps) (batched.py:313-319)"""
    P = bgemm_nt(X, _tr(B), out_exp="col", bt_rowexp=B.expt())
    for _ in range(refine):
        Rr = badd(B, bgemm_nt(Lm, _tr(P), out_exp="col", bt_rowexp=P.expt()), -1, out_exp="align")
        P = badd(P, bgemm_nt(X, _tr(Rr), out_exp="col", bt_rowexp=Rr.expt()), 1, out_exp="align")
    return P


def btrsm_left(Lf, B, tiles, bt=32, refine=0):
    """P = L^{-1} B batched; B (nb, n, m) with per-(block, column) exponents (batched.py:322-335)"""
    nb, n, m = B.shape
    if n <= bt:
        t = tiles[0]
        return _bleaf(t.X, t.Lmp, B, refine)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    P1 = btrsm_left(Lf[:n1, :n1], B[:n1, :], tiles[:k1], bt, refine)
    U = bgemm_nt(Lf[n1:, :n1], _tr(P1), out_exp="col", bt_rowexp=P1.expt())
    B2 = badd(B[n1:, :], U, -1, out_exp="align")
    P2 = btrsm_left(Lf[n1:, n1:], B2, tiles[k1:], bt, refine)
    ee = torch.maximum(P1.expt(), P2.expt())
    P1, P2 = brescale(P1, ee), brescale(P2, ee)
    return BMP(torch.cat([P1.limbs, P2.limbs], 2), ee)


def btrsm_left_T(Lf, B, tiles, bt=32, refine=1):
    """P = L^{-T} B batched (L lower, fixed point from bchol) (batched.py:338-352)"""
    nb, n, m = B.shape
    if n <= bt:
        t = tiles[0]
        XT = BMP(t.X.limbs.transpose(2, 3).contiguous(), t.X.exp)
        return _bleaf(XT, t.Lmp.t(), B, refine)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    P2 = btrsm_left_T(Lf[n1:, n1:], B[n1:, :], tiles[k1:], bt, refine)
    U = bgemm_nt(Lf[n1:, :n1].t(), _tr(P2), out_exp="col", bt_rowexp=P2.expt())
    B1 = badd(B[:n1, :], U, -1, out_exp="align")
    P1 = btrsm_left_T(Lf[:n1, :n1], B1, tiles[:k1], bt, refine)
    ee = torch.maximum(P1.expt(), P2.expt())
    P1, P2 = brescale(P1, ee), brescale(P2, ee)
    return BMP(torch.cat([P1.limbs, P2.limbs], 2), ee)
