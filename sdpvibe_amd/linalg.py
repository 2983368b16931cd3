"""Single-matrix multiprecision Cholesky and triangular solves (the reference's sdpvibe/linalg.py
interface), built from exact RNS GEMMs and the tile base case.

All routines walk the split tree of batched.py (the dimension halved at multiples of the tile size
bt; leaves are diagonal tiles), so the split points and rounding points are the reference's and the
results bit-identical to it:

- chol_inv(A): L and its inverse X, block by block: with the head factored (L11, X11), the panel is
  L21 = A21 X11^T, the tail S = A22 - L21 L21^T (exact at the fixed-point exponent) is factored,
  and X21 = -X22 (L21 X11);
- chol / trsm_right / trsm_left: the stable variant, recursive triangular solves whose leaves use
  either forward substitution on the factored tile ('subst') or the tile's explicit inverse ('gemm');
- trsm(X, B) = X B, one GEMM with per-column exponents.
"""
import torch

from .basecase import apply_tile_inverse, chol_inv_tile, chol_tile, trsv_tile
from .batched import _tree
from .mpmat import MP, add, gemm_nt, neg, rescale


def fx_exp(L):
    """exponent of the equilibrated fixed-point format: |x| <= 1 with 40 bits of headroom below the
    top limb"""
    return -(32 * L - 40)


def _lower_blocks(L11, L21, L22, upper_zero=True):
    """[[L11, 0], [L21, L22]] at the largest of the three (uniform) exponents, the others re-rounded"""
    parts = (L11, L21, L22)
    e = max(p.uniform_exp() for p in parts)
    parts = [p if p.uniform_exp() == e else rescale(p, e) for p in parts]
    k, m = L11.shape[0], L22.shape[0]
    out = torch.zeros((L11.L, k + m, k + m), dtype=torch.int32, device=L11.limbs.device)
    out[:, :k, :k] = parts[0].limbs
    out[:, k:, :k] = parts[1].limbs
    out[:, k:, k:] = parts[2].limbs
    return MP(out, e)


def _identity_padded(A, bt):
    """A (n < bt) in the top-left corner of a bt x bt fixed-point identity"""
    n, L = A.shape[0], A.L
    one = (1 << (32 * L - 40)).to_bytes(4 * L, "little", signed=True)
    words = torch.frombuffer(bytearray(one), dtype=torch.int32).to(A.limbs.device)
    P = torch.zeros((L, bt, bt), dtype=torch.int32, device=A.limbs.device)
    P[:, :n, :n] = A.limbs
    idx = torch.arange(n, bt, device=A.limbs.device)
    P[:, idx, idx] = words[:, None]
    return MP(P, A.uniform_exp())


def chol_inv(A, bt=32, stats=None):
    """(L, X = L^-1) of an SPD equilibrated fixed-point A at fx_exp(L); any n (leaves smaller than bt
    are padded with the identity)"""
    e = fx_exp(A.L)
    if A.uniform_exp() != e:
        raise ValueError("chol_inv takes the fixed-point format fx_exp(L)")

    def walk(node, M):
        if node.head is None:
            if node.n == bt:
                return chol_inv_tile(M)
            Lt, Xt = chol_inv_tile(_identity_padded(M, bt))
            return Lt[:node.n, :node.n], Xt[:node.n, :node.n]
        k = node.k
        L11, X11 = walk(node.head, M[:k, :k])
        L21 = gemm_nt(M[k:, :k], X11, out_exp=e)                      # A21 X11^T
        L22, X22 = walk(node.tail, add(M[k:, k:], gemm_nt(L21, L21, out_exp=e), -1, out_exp=e))
        T = gemm_nt(L21, X11.t(), out_exp="mat")                        # L21 X11
        X21 = neg(gemm_nt(X22, T.t(), out_exp="mat"))                   # -X22 L21 X11
        return _lower_blocks(L11, L21, L22), _lower_blocks(X11, X21, X22)

    return walk(_tree(A.shape[0], bt), A)


def trsm(X, B, kc=4096):
    """X B (X = L^-1 from chol_inv); B may have per-column exponents, and so does the result"""
    rows = MP(B.limbs.transpose(1, 2).contiguous(), 0)
    return gemm_nt(X, rows, out_exp="col", kc=kc, bt_rowexp=B.exp)


def chol(A, bt=32, tile_inverses=False, panel_leaf="subst", refine=0):
    """A (n x n, n a multiple of bt) SPD equilibrated at fx_exp(L) -> (L, tiles in tree order)"""
    L = A.L
    e = fx_exp(L)
    n = A.shape[0]
    out = torch.zeros((L, n, n), dtype=torch.int32, device=A.limbs.device)
    tiles = []

    def factor(node, M):
        lo = node.lo
        if node.head is None:
            if node.n != bt:
                raise ValueError("the dimension must be a multiple of the tile size")
            r = chol_tile(M, want_inverse=tile_inverses)
            tiles.append(r[1])
            out[:, lo:lo + bt, lo:lo + bt] = r[0].limbs
            return
        k, hi = node.k, lo + node.n
        factor(node.head, M[:k, :k])
        L11 = MP(out[:, lo:lo + k, lo:lo + k].contiguous(), e)
        panel = _right(node.head, M[k:, :k], L11, tiles, panel_leaf, refine)
        out[:, lo + k:hi, lo:lo + k] = panel.limbs
        factor(node.tail, add(M[k:, k:], gemm_nt(panel, panel, out_exp=e), -1, out_exp=e))

    factor(_tree(n, bt), A)
    return MP(out, e), tiles


def _right(node, B, Lf, tiles, leaf, refine):
    """X with X Lf^T = B on the node's diagonal block"""
    e = fx_exp(Lf.L)
    if node.head is None:
        t = tiles[node.first]
        if leaf == "gemm" and t.X is not None:
            X = gemm_nt(B, t.X, out_exp=e)                              # B X_t^T
            for _ in range(refine):                                     # one residual correction per step
                resid = add(B, gemm_nt(X, t.Lmp, out_exp=e), -1, out_exp=e)
                X = add(X, gemm_nt(resid, t.X, out_exp=e), 1, out_exp=e)
            return X
        rows = B.shape[0]
        limbs, _ = trsv_tile(t, B.limbs, torch.full((rows,), e, dtype=torch.int64, device=B.limbs.device), out_exp=e)
        return MP(limbs, e)
    k = node.k
    X1 = _right(node.head, B[:, :k], Lf[:k, :k], tiles, leaf, refine)
    rest = add(B[:, k:], gemm_nt(X1, Lf[k:, :k], out_exp=e), -1, out_exp=e)
    X2 = _right(node.tail, rest, Lf[k:, k:], tiles, leaf, refine)
    return MP(torch.cat([X1.limbs, X2.limbs], 2), e)


def trsm_right(B, Lf, tiles, bt=32, leaf="subst", refine=0):
    """X with X L^T = B (B: m x n fixed point; X fixed point, e.g. a Cholesky panel)"""
    return _right(_tree(Lf.shape[0], bt), B, Lf, tiles, leaf, refine)


def _left(node, Lf, B, tiles, leaf, refine):
    if node.head is None:
        t = tiles[node.first]
        if leaf == "gemm" and t.X is not None:
            return apply_tile_inverse(t, B, refine)
        limbs, ex = trsv_tile(t, B.limbs.transpose(1, 2).contiguous(), B.exp)
        return MP(limbs.transpose(1, 2).contiguous(), ex)
    k = node.k
    P1 = _left(node.head, Lf[:k, :k], B[:k, :], tiles, leaf, refine)
    coupled = gemm_nt(Lf[k:, :k], MP(P1.limbs.transpose(1, 2).contiguous(), 0), out_exp="col", bt_rowexp=P1.exp)
    P2 = _left(node.tail, Lf[k:, k:], add(B[k:, :], coupled, -1, out_exp="align"), tiles, leaf, refine)
    e = torch.maximum(P1.exp, P2.exp)                                   # one exponent per column
    return MP(torch.cat([rescale(P1, e).limbs, rescale(P2, e).limbs], 1), e)


def trsm_left(Lf, B, tiles, bt=32, leaf="gemm", refine=0):
    """L^-1 B, B (n x m) with per-column exponents, and so is the result. leaf: 'subst' (substitution
    on the factored tile) or 'gemm' (explicit tile inverses)"""
    return _left(_tree(Lf.shape[0], bt), Lf, B, tiles, leaf, refine)
