"""TEST INFRASTRUCTURE ONLY: ctypes binding of oracle/liboracle.so, the GMP-mpz CPU restatement
of the reference semantics (see oracle.c). Imported by tests/, the smoke check and bench.py's
CPU-baseline leg only, as the checker; the product package never imports it.

All arrays are host numpy arrays in the library's layouts.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
EZ = -(1 << 40)
_P, _I64, _I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
_SIGS = {
    "orc_shift_round": [_P, _I64, _I64, _P, _P, _I64, _P],
    "orc_add_same": [_P, _P, _P, _I64, _I64, _I32],
    "orc_bitlen": [_P, _I64, _I64, _P],
    "orc_badd_align": [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I32],
    "orc_mpf_top": [_P, _P, _I64, _I64, _P],
    "orc_mpf_norm": [_P, _P, _I64, _P, _P, _I64, _I64],
    "orc_mpf_mul": [_P, _P, _P, _P, _P, _P, _I64, _I64, _I32],
    "orc_mpf_add": [_P, _P, _P, _P, _P, _P, _I64, _I64, _I32],
    "orc_residues": [_P, _I64, _I64, _I64, _I64, _I64, _P, _I64, _P],
    "orc_crt": [_P, _P, _I64, _I64, _P, _I64, _P],
    "orc_mpf_gemm": [_P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P, _P, _I64, _I64, _I64,
                     _I64, _P, _P],
    "orc_gemm_exact": [_P, _P, _I64, _I64, _I64, _I64, _P, _I64],
    "orc_bgemm_nt": [_P, _P, _I64, _I64, _I64, _I64, _I64, _P, _I32, _P, _P],
    "orc_tile_chol": [_P, _P, _P, _I64, _I32, _I32, _I32],
    "orc_tile_trsv": [_P, _P, _P, _I64, _I32, _I32, _I32],
}
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        h = ctypes.CDLL(LIB)
        for k, a in _SIGS.items():
            getattr(h, k).argtypes = a
            getattr(h, k).restype = None
        _lib = h
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def shift_round(X, s, Lout):
    X, s = _c(X, np.int32), _c(s, np.int64)
    Lin, B = X.shape
    out = np.zeros((Lout, B), np.int32)
    flag = np.zeros(B, np.int32)
    lib().orc_shift_round(_p(X), Lin, B, _p(s), _p(out), Lout, _p(flag))
    return out, flag


def add_same(A, Bm, sgn=1):
    A, Bm = _c(A, np.int32), _c(Bm, np.int32)
    L, B = A.shape
    out = np.zeros_like(A)
    lib().orc_add_same(_p(A), _p(Bm), _p(out), L, B, sgn)
    return out


def bitlen(X):
    X = _c(X, np.int32)
    L, B = X.shape
    out = np.zeros(B, np.int64)
    lib().orc_bitlen(_p(X), L, B, _p(out))
    return out


def badd_align(A, Bm, EA, EB, sgn):
    """A, Bm: (L, nb, rows, cols); EA, EB: (nb, cols)"""
    A, Bm, EA, EB = _c(A, np.int32), _c(Bm, np.int32), _c(EA, np.int64), _c(EB, np.int64)
    L, nb, rows, cols = A.shape
    out = np.zeros_like(A)
    EO = np.zeros_like(EA)
    lib().orc_badd_align(_p(A), _p(Bm), _p(EA), _p(EB), _p(out), _p(EO), L, nb, rows, cols, sgn)
    return out, EO


def mpf_top(m, e):
    m, e = _c(m, np.int32), _c(e, np.int64)
    L = m.shape[0]
    out = np.zeros(e.size, np.int64)
    lib().orc_mpf_top(_p(m), _p(e), L, e.size, _p(out))
    return out.reshape(e.shape)


def mpf_norm(X, eI, L):
    X, eI = _c(X, np.int32), _c(eI, np.int64)
    LI, n = X.shape
    O = np.zeros((L, n), np.int32)
    eO = np.zeros(n, np.int64)
    lib().orc_mpf_norm(_p(X), _p(eI), LI, _p(O), _p(eO), L, n)
    return O, eO


def mpf_mul(A, eA, Bm, eB):
    A, eA, Bm, eB = _c(A, np.int32), _c(eA, np.int64), _c(Bm, np.int32), _c(eB, np.int64)
    L = A.shape[0]
    n = eA.size
    bs = int(eB.size == 1)
    O = np.zeros((L, n), np.int32)
    eO = np.zeros(n, np.int64)
    lib().orc_mpf_mul(_p(A), _p(eA), _p(Bm), _p(eB), _p(O), _p(eO), L, n, bs)
    return O, eO


def mpf_add(A, eA, Bm, eB, sgn=1):
    A, eA, Bm, eB = _c(A, np.int32), _c(eA, np.int64), _c(Bm, np.int32), _c(eB, np.int64)
    L = A.shape[0]
    n = eA.size
    O = np.zeros((L, n), np.int32)
    eO = np.zeros(n, np.int64)
    lib().orc_mpf_add(_p(A), _p(eA), _p(Bm), _p(eB), _p(O), _p(eO), L, n, sgn)
    return O, eO


def residues(X, k0, kc, mods):
    """X (L, rows, K) -> centred residues (G, rows, kc) int32"""
    X, mods = _c(X, np.int32), _c(mods, np.int32)
    L, rows, K = X.shape
    out = np.zeros((mods.size, rows, kc), np.int32)
    lib().orc_residues(_p(X), L, rows, K, k0, kc, _p(mods), mods.size, _p(out))
    return out


def crt(acc, mods, nw):
    """acc (n, B) residues -> (words (nw, B) int32, bitlen (B,))"""
    acc, mods = _c(acc, np.int16), _c(mods, np.int32)
    n, B = acc.shape
    W = np.zeros((nw, B), np.int32)
    bl = np.zeros(B, np.int64)
    lib().orc_crt(_p(acc), _p(mods), n, B, _p(W), nw, _p(bl))
    return W, bl


def _geom(shape, trans):
    _, r, c = shape
    return (c, r, r * c, 1, c) if trans else (r, c, r * c, c, 1)


def mpf_gemm(Am, Ae, Bm, Be, transA=False, transB=False):
    """mpf.gemm of batched (nb, ., .) operands: m (L, nb, r, c), e (nb, r, c) -> (m, e)"""
    Am, Ae, Bm, Be = _c(Am, np.int32), _c(Ae, np.int64), _c(Bm, np.int32), _c(Be, np.int64)
    L = Am.shape[0]
    nb = Ae.shape[0]
    M, K, aldb, aldr, aldk = _geom(Ae.shape, transA)
    N, K2, bldb, bldr, bldk = _geom(Be.shape, not transB)
    assert K == K2
    out = np.zeros((L, nb, M, N), np.int32)
    oe = np.zeros((nb, M, N), np.int64)
    lib().orc_mpf_gemm(_p(Am), _p(Ae), L, nb, M, N, K, aldb, aldr, aldk, Ae.size, _p(Bm), _p(Be), bldb, bldr,
                       bldk, Be.size, _p(out), _p(oe))
    return out, oe


def gemm_exact(A, Bt, nw):
    """A (L, M, K), Bt (L, N, K) limb planes -> exact A Bt^T as nw words (nw, M, N)"""
    A, Bt = _c(A, np.int32), _c(Bt, np.int32)
    L, M, K = A.shape
    N = Bt.shape[1]
    W = np.zeros((nw, M, N), np.int32)
    lib().orc_gemm_exact(_p(A), _p(Bt), L, M, N, K, _p(W), nw)
    return W


def bgemm_nt(A, Bt, e_in, mode, oexp=None):
    """batched.bgemm_nt exact product + _bcrt_round. A (L, nb, M, K), Bt (L, nb, N, K) limb planes,
    e_in (nb, N) exponents of the exact product; mode 0 fixed (oexp (nb, N) given), 1 'col', 2 'mat'.
    Returns (out (L, nb, M, N), oexp (nb, N))."""
    A, Bt = _c(A, np.int32), _c(Bt, np.int32)
    L, nb, M, K = A.shape
    N = Bt.shape[2]
    e_in = _c(np.broadcast_to(e_in, (nb, N)), np.int64)
    oe = np.zeros((nb, N), np.int64) if oexp is None else _c(np.broadcast_to(oexp, (nb, N)), np.int64).copy()
    out = np.zeros((L, nb, M, N), np.int32)
    lib().orc_bgemm_nt(_p(A), _p(Bt), L, nb, M, N, K, _p(e_in), mode, _p(out), _p(oe))
    return out, oe


def tile_chol(W, LF):
    """chol_tiles + clearing + trinv_tiles (basecase.py:159-226) on W (nb, BT, BT, WL) uint32 wide
    fixed-point words (F = 32*LF fractional bits). Returns (L (nb, BT, BT, WL), RINV (nb, BT, WL),
    X (nb, BT, BT, WL))."""
    W = np.ascontiguousarray(W, dtype=np.uint32).copy()
    nb, BT, _, WL = W.shape
    RINV = np.zeros((nb, BT, WL), np.uint32)
    X = np.zeros_like(W)
    lib().orc_tile_chol(_p(W), _p(RINV), _p(X), nb, BT, WL, LF)
    return W, RINV, X


def tile_trsv(Lw, RINV, V, LF):
    """trsv_warp (basecase.py:228-245): Lw (BT, BT, WL), RINV (BT, WL), V (nvec, BT, WL) uint32 words;
    returns y = L^-1 v per vector (same layout)"""
    Lw, RINV = np.ascontiguousarray(Lw, dtype=np.uint32), np.ascontiguousarray(RINV, dtype=np.uint32)
    V = np.ascontiguousarray(V, dtype=np.uint32).copy()
    nvec, BT, WL = V.shape
    lib().orc_tile_trsv(_p(Lw), _p(RINV), _p(V), nvec, BT, WL, LF)
    return V
