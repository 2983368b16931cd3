"""TEST INFRASTRUCTURE ONLY: the batched recursive Cholesky and triangular solves
(sdpvibe/batched.py:243-352) and the equilibrated Cholesky driver (sdpvibe/mpf.py:461-532) restated
over the oracle's primitives (shift_round, add_same, badd_align, bgemm_nt, tile_chol) on host numpy
arrays. The recursion fixes where every rounding happens, so a bit-exact comparison with the GPU
path checks both the kernels and the host orchestration. Imported by tests/ only.
"""
import numpy as np

from . import oracle as O

EZ = O.EZ


def headroom_bits(L):
    return 32 * L - 32


def fx_exp(L):
    return -(32 * L - 40)


class OB:
    """nb matrices: limbs (L, nb, R, C) int32; exp a Python int or an (nb, C) int64 array"""

    def __init__(self, limbs, exp):
        self.limbs = np.ascontiguousarray(limbs, dtype=np.int32)
        self.exp = exp if isinstance(exp, int) else np.ascontiguousarray(exp, dtype=np.int64)

    @property
    def shape(self):
        return self.limbs.shape[1:]

    @property
    def L(self):
        return self.limbs.shape[0]

    def expt(self):
        nb, R, C = self.shape
        if isinstance(self.exp, int):
            return np.full((nb, C), self.exp, np.int64)
        return self.exp

    def sub(self, r, c):
        e = self.exp if isinstance(self.exp, int) else self.exp[:, c]
        return OB(self.limbs[:, :, r, c], e)

    def t(self):
        assert isinstance(self.exp, int)
        return OB(self.limbs.transpose(0, 1, 3, 2), self.exp)


def brescale(A, new_exp, Lout=None):
    """batched.py:45-52"""
    nb, R, C = A.shape
    Lo = Lout or A.L
    if isinstance(new_exp, int) and isinstance(A.exp, int):
        s = np.full(nb * R * C, new_exp - A.exp, np.int64)
        out, _ = O.shift_round(A.limbs.reshape(A.L, -1), s, Lo)
        return OB(out.reshape(Lo, nb, R, C), new_exp)
    ne = np.broadcast_to(np.asarray(new_exp, np.int64), (nb, C))
    s = np.broadcast_to((ne - A.expt())[:, None, :], (nb, R, C)).reshape(-1)
    out, _ = O.shift_round(A.limbs.reshape(A.L, -1), s, Lo)
    return OB(out.reshape(Lo, nb, R, C), ne.copy())


def badd(A, B, alpha, out_exp):
    """batched.py:82-92: exact at a common int exponent, else 'align'"""
    L = A.L
    if isinstance(out_exp, int) and A.exp == out_exp and B.exp == out_exp:
        return OB(O.add_same(A.limbs.reshape(L, -1), B.limbs.reshape(L, -1), alpha).reshape(A.limbs.shape),
                  out_exp)
    assert out_exp == "align"
    out, eo = O.badd_align(A.limbs, B.limbs, A.expt(), B.expt(), alpha)
    return OB(out, eo)


def bgemm_nt(A, Bt, out_exp="col", bt_rowexp=None):
    """batched.py:165-231"""
    nb, M, K = A.shape
    N = Bt.shape[1]
    eA = A.exp if isinstance(A.exp, int) else A.exp[:, 0]
    eB = Bt.exp if bt_rowexp is None else bt_rowexp
    eA_t = np.broadcast_to(np.asarray(eA, np.int64).reshape(-1, 1) if not isinstance(eA, int) else eA, (nb, N))
    eB_t = np.asarray(eB, np.int64)
    if eB_t.ndim == 1:
        eB_t = eB_t[:, None]
    e_in = np.broadcast_to(eA_t + eB_t, (nb, N))
    if isinstance(out_exp, int):
        out, _ = O.bgemm_nt(A.limbs, Bt.limbs, e_in, 0, np.full((nb, N), out_exp, np.int64))
        return OB(out, out_exp)
    out, oe = O.bgemm_nt(A.limbs, Bt.limbs, e_in, 1 if out_exp == "col" else 2)
    return OB(out, oe)


def _soa_to_aos(X):
    """(WL, nb, b, b) int32 planes -> (nb, b, b, WL) uint32 words"""
    return np.ascontiguousarray(X.transpose(1, 2, 3, 0)).view(np.uint32)


def _aos_to_soa(W):
    return np.ascontiguousarray(W.view(np.int32).transpose(3, 0, 1, 2))


class Tile:
    __slots__ = ("X", "Lmp")


def bchol_tile(A, LI=24):
    """batched.py:243-262"""
    nb, b, _ = A.shape
    L, e = A.L, A.exp
    WL, F = L + LI, 32 * L
    Ws = brescale(A, -F, WL).limbs
    Lw, _, X = O.tile_chol(_soa_to_aos(Ws), L)
    Lout = brescale(OB(_aos_to_soa(Lw), -F), e, L)
    Xs = _aos_to_soa(X)
    top = O.bitlen(Xs.reshape(WL, -1)).reshape(nb, b * b).max(1) - F
    ex = np.broadcast_to((top - headroom_bits(L))[:, None], (nb, b)).copy()
    t = Tile()
    t.Lmp = Lout
    t.X = brescale(OB(Xs, np.full((nb, b), -F, np.int64)), ex, L)
    return Lout, t


def _split(n, bt):
    return (n // 2 + bt - 1) // bt * bt


def _ntiles(n, bt):
    if n <= bt:
        return 1
    n1 = _split(n, bt)
    return _ntiles(n1, bt) + _ntiles(n - n1, bt)


def bchol(A, bt=32, refine=1):
    """batched.py:265-288"""
    nb, n, _ = A.shape
    L, e = A.L, A.exp
    if n <= bt:
        Lt, t = bchol_tile(A)
        return Lt, [t]
    n1 = _split(n, bt)
    L11, t1 = bchol(A.sub(slice(0, n1), slice(0, n1)), bt, refine)
    L21 = btrsm_right(A.sub(slice(n1, n), slice(0, n1)), L11, t1, bt, refine)
    S = badd(A.sub(slice(n1, n), slice(n1, n)), bgemm_nt(L21, L21, out_exp=e), -1, out_exp=e)
    L22, t2 = bchol(S, bt, refine)
    out = np.zeros((L, nb, n, n), np.int32)
    out[:, :, :n1, :n1] = L11.limbs
    out[:, :, n1:, :n1] = L21.limbs
    out[:, :, n1:, n1:] = L22.limbs
    return OB(out, e), t1 + t2


def btrsm_right(B, Lf, tiles, bt=32, refine=1):
    """batched.py:291-305"""
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
    X1 = btrsm_right(B.sub(slice(None), slice(0, n1)), Lf.sub(slice(0, n1), slice(0, n1)), tiles[:k1], bt, refine)
    B2 = badd(B.sub(slice(None), slice(n1, n)), bgemm_nt(X1, Lf.sub(slice(n1, n), slice(0, n1)), out_exp=e), -1,
              out_exp=e)
    X2 = btrsm_right(B2, Lf.sub(slice(n1, n), slice(n1, n)), tiles[k1:], bt, refine)
    return OB(np.concatenate([X1.limbs, X2.limbs], 3), e)


def _tr(B):
    return OB(B.limbs.transpose(0, 1, 3, 2), 0)


def _bleaf(X, Lm, B, refine):
    """batched.py:313-319"""
    P = bgemm_nt(X, _tr(B), out_exp="col", bt_rowexp=B.expt())
    for _ in range(refine):
        Rr = badd(B, bgemm_nt(Lm, _tr(P), out_exp="col", bt_rowexp=P.expt()), -1, out_exp="align")
        P = badd(P, bgemm_nt(X, _tr(Rr), out_exp="col", bt_rowexp=Rr.expt()), 1, out_exp="align")
    return P


def _join_rows(P1, P2):
    ee = np.maximum(P1.expt(), P2.expt())
    P1, P2 = brescale(P1, ee), brescale(P2, ee)
    return OB(np.concatenate([P1.limbs, P2.limbs], 2), ee)


def btrsm_left(Lf, B, tiles, bt=32, refine=0):
    """batched.py:322-335"""
    nb, n, m = B.shape
    if n <= bt:
        t = tiles[0]
        return _bleaf(t.X, t.Lmp, B, refine)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    P1 = btrsm_left(Lf.sub(slice(0, n1), slice(0, n1)), B.sub(slice(0, n1), slice(None)), tiles[:k1], bt, refine)
    U = bgemm_nt(Lf.sub(slice(n1, n), slice(0, n1)), _tr(P1), out_exp="col", bt_rowexp=P1.expt())
    B2 = badd(B.sub(slice(n1, n), slice(None)), U, -1, out_exp="align")
    P2 = btrsm_left(Lf.sub(slice(n1, n), slice(n1, n)), B2, tiles[k1:], bt, refine)
    return _join_rows(P1, P2)


def btrsm_left_T(Lf, B, tiles, bt=32, refine=1):
    """batched.py:338-352"""
    nb, n, m = B.shape
    if n <= bt:
        t = tiles[0]
        XT = OB(t.X.limbs.transpose(0, 1, 3, 2), t.X.exp)
        return _bleaf(XT, t.Lmp.t(), B, refine)
    n1 = _split(n, bt)
    k1 = _ntiles(n1, bt)
    P2 = btrsm_left_T(Lf.sub(slice(n1, n), slice(n1, n)), B.sub(slice(n1, n), slice(None)), tiles[k1:], bt, refine)
    U = bgemm_nt(Lf.sub(slice(n1, n), slice(0, n1)).t(), _tr(P2), out_exp="col", bt_rowexp=P2.expt())
    B1 = badd(B.sub(slice(0, n1), slice(None)), U, -1, out_exp="align")
    P1 = btrsm_left_T(Lf.sub(slice(0, n1), slice(0, n1)), B1, tiles[:k1], bt, refine)
    return _join_rows(P1, P2)


# ------------------------------------------------------------------ equilibrated Cholesky driver
def align_rows(m, e):
    """mpf.py:390-398 on host arrays: m (L, ..., R, K), e (..., R, K) -> (aligned limbs, row exps)"""
    L = m.shape[0]
    NB = 32 * L - 2
    tp = O.mpf_top(m.reshape(L, -1), e.reshape(-1)).reshape(e.shape)
    E = tp.max(-1) - NB
    E = np.where(E < EZ // 2, 0, E)
    s = (E[..., None] - e).reshape(-1)
    s = np.where(e.reshape(-1) == EZ, 1 << 30, s)
    out, _ = O.shift_round(m.reshape(L, -1), s, L)
    return out.reshape(m.shape), E


class Chol:
    """mpf.py:461-532 on host arrays: m (L, nb, n, n), e (nb, n, n)"""

    def __init__(self, m, e, bt=32, refine=1, solve_refine=0):
        L, nb, n, _ = m.shape
        self.L, self.nb, self.n, self.bt, self.solve_refine = L, nb, n, bt, solve_refine
        d = np.arange(n)
        tpd = O.mpf_top(m.reshape(L, -1), e.reshape(-1)).reshape(e.shape)[:, d, d]
        if (m[-1][:, d, d] < 0).any() or (tpd <= EZ // 2).any():
            raise ValueError("matrix is not positive definite (diagonal)")
        k = (tpd + 1) // 2
        self.k = k
        npad = max(bt, (n + bt - 1) // bt * bt)
        self.npad = npad
        efx = fx_exp(L)
        s = efx - (e - k[:, :, None] - k[:, None, :])
        s = np.where(e == EZ, 1 << 30, s).reshape(-1)
        Af, _ = O.shift_round(m.reshape(L, -1), s, L)
        P = np.zeros((L, nb, npad, npad), np.int32)
        P[:, :, :n, :n] = Af.reshape(L, nb, n, n)
        one = np.frombuffer((1 << -efx).to_bytes(4 * L, "little"), dtype=np.int32)    # 1 at exponent efx
        for i in range(n, npad):
            P[:, :, i, i] = one[:, None]
        self.Lt, self.tiles = bchol(OB(P, efx), bt, refine)

    def _to_ob(self, bm, be):
        L = self.L
        Bm, cE = align_rows(np.ascontiguousarray(bm.transpose(0, 1, 3, 2)), np.ascontiguousarray(be.transpose(0, 2, 1)))
        P = np.zeros((L, self.nb, self.npad, bm.shape[-1]), np.int32)
        P[:, :, :self.n] = Bm.transpose(0, 1, 3, 2)
        return OB(P, cE)

    def _from_ob(self, P):
        L, m = self.L, P.shape[2]
        X = np.ascontiguousarray(P.limbs[:, :, :self.n]).reshape(L, -1)
        ein = np.broadcast_to(P.expt()[:, None, :], (self.nb, self.n, m)).reshape(-1)
        om, oe = O.mpf_norm(X, ein, L)
        return om.reshape(L, self.nb, self.n, m), oe.reshape(self.nb, self.n, m)

    def solve_lower(self, bm, be):
        be = np.where(be == EZ, EZ, be - self.k[:, :, None])
        return self._from_ob(btrsm_left(self.Lt, self._to_ob(bm, be), self.tiles, self.bt, self.solve_refine))

    def solve_upper(self, bm, be):
        om, oe = self._from_ob(btrsm_left_T(self.Lt, self._to_ob(bm, be), self.tiles, self.bt, self.solve_refine))
        return om, np.where(oe == EZ, EZ, oe - self.k[:, :, None])
