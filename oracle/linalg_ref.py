"""TEST INFRASTRUCTURE ONLY: the single-matrix Cholesky / triangular solves (sdpvibe/linalg.py:19-138),
the tile API (sdpvibe/basecase.py:259-332) and the fixed-point matrix operations they use
(sdpvibe/mpmat.py:108-236) restated on host arrays over the oracle's primitives. The recursion fixes
every rounding point, so the GPU path is compared with it bit for bit. Imported by tests/ only."""
import numpy as np

from . import oracle as O

EZ = O.EZ
LI = 24


def headroom_bits(L):
    return 32 * L - 32


def fx_exp(L):
    return -(32 * L - 40)


class OM:
    """limbs (L, R, C) int32; exp a Python int (uniform) or a (C,) int64 array"""

    def __init__(self, limbs, exp):
        self.limbs = np.ascontiguousarray(limbs, dtype=np.int32)
        self.exp = exp if isinstance(exp, (int, np.integer)) and not isinstance(exp, np.ndarray) else \
            np.ascontiguousarray(np.broadcast_to(np.asarray(exp, np.int64), (self.limbs.shape[2],)))
        if isinstance(self.exp, np.integer):
            self.exp = int(self.exp)

    @property
    def shape(self):
        return self.limbs.shape[1:]

    @property
    def L(self):
        return self.limbs.shape[0]

    def expv(self):
        return np.full(self.shape[1], self.exp, np.int64) if isinstance(self.exp, int) else self.exp

    def uniform_exp(self):
        if isinstance(self.exp, int):
            return self.exp
        assert (self.exp == self.exp[0]).all()
        return int(self.exp[0])

    def t(self):
        return OM(self.limbs.transpose(0, 2, 1), self.uniform_exp())

    def __getitem__(self, idx):
        r, c = idx
        e = self.exp if isinstance(self.exp, int) else self.exp[c]
        return OM(self.limbs[:, r, c], e)


def rescale(A, new_exp, Lout=None):
    """mpmat.py:108-117"""
    R, C = A.shape
    Lo = Lout or A.L
    if isinstance(new_exp, int) and isinstance(A.exp, int):
        out, _ = O.shift_round(A.limbs.reshape(A.L, -1), np.full(R * C, new_exp - A.exp, np.int64), Lo)
        return OM(out.reshape(Lo, R, C), new_exp)
    ne = np.broadcast_to(np.asarray(new_exp, np.int64), (C,)).copy()
    s = np.broadcast_to((ne - A.expv())[None, :], (R, C)).reshape(-1)
    out, _ = O.shift_round(A.limbs.reshape(A.L, -1), s, Lo)
    return OM(out.reshape(Lo, R, C), ne)


def neg(A):
    z = np.zeros_like(A.limbs.reshape(A.L, -1))
    return OM(O.add_same(z, A.limbs.reshape(A.L, -1), -1).reshape(A.limbs.shape), A.exp)


def _bitlen(limbs):
    return O.bitlen(limbs.reshape(limbs.shape[0], -1)).reshape(limbs.shape[1:])


def add(A, B, alpha=1, out_exp=None):
    """mpmat.py:135-156"""
    R, C = A.shape
    L = A.L
    if isinstance(out_exp, int) and isinstance(A.exp, int) and isinstance(B.exp, int) and A.exp == out_exp \
            and B.exp == out_exp:
        return OM(O.add_same(A.limbs.reshape(L, -1), B.limbs.reshape(L, -1), alpha).reshape(L, R, C), out_exp)
    if out_exp == "align":
        e = np.maximum(A.expv(), B.expv()) + 1
        a, b = rescale(A, e), rescale(B, e)
        return OM(O.add_same(a.limbs.reshape(L, -1), b.limbs.reshape(L, -1), alpha).reshape(L, R, C), e)
    e = np.minimum(A.expv(), B.expv())
    d = int(max((A.expv() - e).max(), (B.expv() - e).max()))
    Lw = L + (d + 31) // 32 + 2
    a, b = rescale(A, e, Lw), rescale(B, e, Lw)
    W = OM(O.add_same(a.limbs.reshape(Lw, -1), b.limbs.reshape(Lw, -1), alpha).reshape(Lw, R, C), e.copy())
    if out_exp is None or out_exp in ("mat", "col"):
        top = _bitlen(W.limbs) + W.expv()[None, :]
        top = top.max(0) if out_exp == "col" else np.full(C, top.max())
        out_exp = top - headroom_bits(L)
    return rescale(W, out_exp, L)


def gemm_nt(A, Bt, out_exp="col", bt_rowexp=None):
    """mpmat.py:169-236 (one block of the batched restatement)"""
    M, K = A.shape
    N = Bt.shape[0]
    eA = A.uniform_exp()
    e_in = eA + (np.asarray(bt_rowexp, np.int64) if bt_rowexp is not None else Bt.uniform_exp())
    e_in = np.broadcast_to(np.asarray(e_in, np.int64), (N,))[None, :]
    Al, Bl = A.limbs[:, None], Bt.limbs[:, None]
    if isinstance(out_exp, str):
        out, oe = O.bgemm_nt(Al, Bl, e_in, 1)
        C = OM(out[:, 0], oe[0])
        if out_exp == "mat":
            C = rescale(C, int(oe.max()))
        return C
    oe = np.broadcast_to(np.asarray(out_exp, np.int64), (N,))[None, :]
    out, _ = O.bgemm_nt(Al, Bl, e_in, 0, oe)
    uniform = isinstance(out_exp, (int, np.integer)) and not isinstance(out_exp, np.ndarray)
    fixed_int = uniform and bt_rowexp is None
    return OM(out[:, 0], int(out_exp) if fixed_int else oe[0].copy())


# ------------------------------------------------------------------ tile API (basecase.py:259-332)
class Tile:
    __slots__ = ("Lw", "RINV", "b", "WL", "LF", "X", "Lmp")


def _soa_to_aos(X):
    return np.ascontiguousarray(np.moveaxis(X, 0, -1)).view(np.uint32)


def _aos_to_soa(W):
    return np.ascontiguousarray(np.moveaxis(W.view(np.int32), -1, 0))


def chol_tile(A, want_inverse=False):
    b = A.shape[0]
    L, e = A.L, A.uniform_exp()
    WL, F = L + LI, 32 * L
    W = _soa_to_aos(rescale(A, -F, WL).limbs)[None]               # (1, b, b, WL)
    Lw, RI, X = O.tile_chol(W, L)
    Lout = rescale(OM(_aos_to_soa(Lw[0]), -F), e, L)
    t = Tile()
    t.Lw, t.RINV, t.b, t.WL, t.LF = Lw[0], RI[0], b, WL, L
    t.X, t.Lmp = None, Lout
    if not want_inverse:
        return Lout, t
    Xs = _aos_to_soa(X[0])
    top = int(_bitlen(Xs).max()) - F
    t.X = rescale(OM(Xs, -F), top - headroom_bits(L), L)
    return Lout, t, t.X


def chol_inv_tile(A):
    Lout, _, X = chol_tile(A, want_inverse=True)
    return Lout, X


def apply_tile_inverse(t, B, refine=0):
    def tr(M):
        return OM(M.limbs.transpose(0, 2, 1), 0)
    P = gemm_nt(t.X, tr(B), out_exp="col", bt_rowexp=B.expv())
    for _ in range(refine):
        LP = gemm_nt(t.Lmp, tr(P), out_exp="col", bt_rowexp=P.expv())
        Rr = add(B, LP, -1, out_exp="col")
        D = gemm_nt(t.X, tr(Rr), out_exp="col", bt_rowexp=Rr.expv())
        P = add(P, D, 1, out_exp="align")
    return P


def trsv_tile(t, Vlimbs, vexp, out_exp=None):
    L, nvec, b = Vlimbs.shape
    F, FX = 32 * t.LF, 32 * L - 40
    n = nvec * b
    Ws, _ = O.shift_round(Vlimbs.reshape(L, n), np.full(n, FX - F, np.int64), t.WL)
    W = _soa_to_aos(Ws.reshape(t.WL, nvec, b))
    Y = O.tile_trsv(t.Lw, t.RINV, W, t.LF)
    Ws = _aos_to_soa(Y).reshape(t.WL, n)
    base = np.asarray(vexp, np.int64) + FX - F
    if out_exp is None:
        ne = O.bitlen(Ws).reshape(nvec, b).max(1) + base - headroom_bits(L)
    else:
        ne = np.full(nvec, int(out_exp), np.int64)
    s = np.broadcast_to((ne - base)[:, None], (nvec, b)).reshape(-1)
    out, _ = O.shift_round(Ws, s, L)
    return out.reshape(L, nvec, b), ne


# ------------------------------------------------------------------ linalg.py:19-138
def _cat(blocks, L):
    e = max(b.uniform_exp() for row in blocks for b in row if b is not None)
    n = [next(b for b in blocks[i] if b is not None).shape[0] for i in (0, 1)]
    m = [next(blocks[i][j] for i in (0, 1) if blocks[i][j] is not None).shape[1] for j in (0, 1)]
    out = np.zeros((L, n[0] + n[1], m[0] + m[1]), np.int32)
    for i, r0 in ((0, 0), (1, n[0])):
        for j, c0 in ((0, 0), (1, m[0])):
            b = blocks[i][j]
            if b is None:
                continue
            if b.uniform_exp() != e:
                b = rescale(b, e)
            out[:, r0:r0 + n[i], c0:c0 + m[j]] = b.limbs
    return OM(out, e)


def _split(n, bt):
    return (n // 2 + bt - 1) // bt * bt


def _ntiles(n, bt):
    if n <= bt:
        return 1
    n1 = _split(n, bt)
    return _ntiles(n1, bt) + _ntiles(n - n1, bt)


def chol_inv(A, bt=32):
    n = A.shape[0]
    L = A.L
    e = fx_exp(L)
    if n <= bt:
        if n < bt:
            P = np.zeros((L, bt, bt), np.int32)
            P[:, :n, :n] = A.limbs
            one = np.frombuffer((1 << (32 * L - 40)).to_bytes(4 * L, "little"), dtype=np.int32)
            for i in range(n, bt):
                P[:, i, i] = one
            Lt, Xt = chol_inv_tile(OM(P, e))
            return Lt[:n, :n], Xt[:n, :n]
        return chol_inv_tile(A)
    n1 = _split(n, bt)
    L11, X11 = chol_inv(A[:n1, :n1], bt)
    L21 = gemm_nt(A[n1:, :n1], X11, out_exp=e)
    S = add(A[n1:, n1:], gemm_nt(L21, L21, out_exp=e), -1, out_exp=e)
    L22, X22 = chol_inv(S, bt)
    T = gemm_nt(L21, X11.t(), out_exp="mat")
    X21 = neg(gemm_nt(X22, T.t(), out_exp="mat"))
    return _cat([[L11, None], [L21, L22]], L), _cat([[X11, None], [X21, X22]], L)


def trsm(X, B):
    return gemm_nt(X, OM(B.limbs.transpose(0, 2, 1), 0), out_exp="col", bt_rowexp=B.expv())


def chol(A, bt=32, tile_inverses=False, panel_leaf="subst", refine=0):
    n = A.shape[0]
    L = A.L
    e = fx_exp(L)
    if n <= bt:
        r = chol_tile(A, want_inverse=tile_inverses)
        return r[0], [r[1]]
    n1 = _split(n, bt)
    L11, t1 = chol(A[:n1, :n1], bt, tile_inverses, panel_leaf, refine)
    L21 = trsm_right(A[n1:, :n1], L11, t1, bt, panel_leaf, refine)
    S = add(A[n1:, n1:], gemm_nt(L21, L21, out_exp=e), -1, out_exp=e)
    L22, t2 = chol(S, bt, tile_inverses, panel_leaf, refine)
    return _cat([[L11, None], [L21, L22]], L), t1 + t2


def trsm_right(B, Lf, tiles, bt=32, leaf="subst", refine=0):
    n = Lf.shape[0]
    e = fx_exp(Lf.L)
    if n <= bt and leaf == "gemm" and tiles[0].X is not None:
        t = tiles[0]
        X = gemm_nt(B, t.X, out_exp=e)
        for _ in range(refine):
            Rr = add(B, gemm_nt(X, t.Lmp, out_exp=e), -1, out_exp=e)
            X = add(X, gemm_nt(Rr, t.X, out_exp=e), 1, out_exp=e)
        return X
    if n <= bt:
        m = B.shape[0]
        out, _ = trsv_tile(tiles[0], B.limbs, np.full(m, e, np.int64), out_exp=e)
        return OM(out, e)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    X1 = trsm_right(B[:, :n1], Lf[:n1, :n1], tiles[:k1], bt, leaf, refine)
    B2 = add(B[:, n1:], gemm_nt(X1, Lf[n1:, :n1], out_exp=e), -1, out_exp=e)
    X2 = trsm_right(B2, Lf[n1:, n1:], tiles[k1:], bt, leaf, refine)
    return OM(np.concatenate([X1.limbs, X2.limbs], 2), e)


def trsm_left(Lf, B, tiles, bt=32, leaf="gemm", refine=0):
    n = Lf.shape[0]
    if n <= bt:
        if leaf == "gemm" and tiles[0].X is not None:
            return apply_tile_inverse(tiles[0], B, refine)
        out, ex = trsv_tile(tiles[0], B.limbs.transpose(0, 2, 1), B.expv())
        return OM(out.transpose(0, 2, 1), ex)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    P1 = trsm_left(Lf[:n1, :n1], B[:n1, :], tiles[:k1], bt, leaf, refine)
    U = gemm_nt(Lf[n1:, :n1], OM(P1.limbs.transpose(0, 2, 1), 0), out_exp="col", bt_rowexp=P1.expv())
    B2 = add(B[n1:, :], U, -1, out_exp="align")
    P2 = trsm_left(Lf[n1:, n1:], B2, tiles[k1:], bt, leaf, refine)
    e = np.maximum(P1.expv(), P2.expv())
    P1, P2 = rescale(P1, e), rescale(P2, e)
    return OM(np.concatenate([P1.limbs, P2.limbs], 1), e)
