"""Base case on the GPU: multiprecision Cholesky + triangular inverse of small tiles and forward
substitution with a factored tile (mirror of sdpvibe/basecase.py).

Wide fixed-point format inside the kernels: WL = L + 24 32-bit words, two's complement, F = 32L
fractional bits. Input tiles are SPD and equilibrated, so |a_ij| <= 1 and |l_ij| <= 1; pivots may
be tiny and their reciprocals (and X = L^-1) large, hence the integer words.
"""
import torch

from ._lib import call, ptr, stream
from .limbops import _errbuf, bitlen_words, shift_round
from .mpmat import MP, headroom_bits, rescale

LI = 24   # integer words of the wide format


def tile_factor(limbs, BT):
    """limbs: (L, nb, BT, BT) fixed point at exponent -(32L-40) -> (Lout (L, nb, BT, BT) at the same
    exponent, upper triangle zero; Xw (L+24, nb, BT, BT) = L^-1 at exponent -32L; work, whose
    first nb*BT*BT*(L+24) words are the wide factor and whose words from 3*nb*BT*BT*(L+24) on are
    the wide pivot reciprocals)"""
    L, nb = limbs.shape[0], limbs.shape[1]
    WL = L + LI
    A = limbs.contiguous()
    Lout = torch.empty_like(A)
    Xw = torch.empty((WL, nb, BT, BT), dtype=torch.int32, device=A.device)
    work = torch.empty((nb * BT * (3 * BT + 1) * WL,), dtype=torch.int32, device=A.device)
    call("sdpv_tile_chol", ptr(A), L, nb, BT, ptr(Lout), ptr(Xw), ptr(work), ptr(_errbuf()), stream())
    return Lout, Xw, work


class Tile:
    """Cholesky factor of a diagonal tile in the kernels' wide, entry-major format (+ pivot
    reciprocals), with the rounded factor Lmp and, when asked for, the explicit inverse X
    (basecase.py:259-261)"""
    __slots__ = ("Lw", "RINV", "b", "WL", "LF", "X", "Lmp")


def chol_tile(A, LI=LI, threads=256, want_inverse=False):
    """A: MP (b x b) fixed point at exponent -(32L-40) -> (Lfac MP with A's exponent, Tile[, X])
    (basecase.py:268-290). The factor and the inverse come from one launch."""
    b = A.shape[0]
    L = A.L
    e = A.uniform_exp()
    if e != -(32 * L - 40):
        raise ValueError("tiles enter at the fixed-point exponent -(32L - 40)")
    WL, F = L + LI, 32 * L
    Lout, Xw, work = tile_factor(A.limbs.reshape(L, 1, b, b), b)
    t = Tile()
    t.Lw = work[:b * b * WL].view(b, b, WL)
    t.RINV = work[3 * b * b * WL:3 * b * b * WL + b * WL].view(b, WL)
    t.b, t.WL, t.LF = b, WL, L
    t.Lmp = MP(Lout.view(L, b, b), e)
    t.X = None
    if not want_inverse:
        return t.Lmp, t
    Xs = Xw.view(WL, b, b)
    top = int(bitlen_words(Xs.reshape(WL, -1)).max()) - F
    t.X = rescale(MP(Xs, -F), top - headroom_bits(L), L)
    return t.Lmp, t, t.X


def chol_inv_tile(A, LI=LI, threads=256):
    """(basecase.py:293-295)"""
    Lout, _, X = chol_tile(A, LI, threads, want_inverse=True)
    return Lout, X


def _rows_of(M):
    """M^T as the row operand of gemm_nt (its column exponents are passed as row exponents)"""
    return MP(M.limbs.transpose(1, 2).contiguous(), 0)


def apply_tile_inverse(t, B, refine=0):
    """L_tile^-1 B through the tile's explicit inverse, then `refine` residual corrections
    (B - L P, multiplied by the inverse, added back aligned); B: MP (b x m) with per-column
    exponents, and so is the result"""
    from .mpmat import add, gemm_nt
    P = gemm_nt(t.X, _rows_of(B), out_exp="col", bt_rowexp=B.exp)
    for _ in range(refine):
        resid = add(B, gemm_nt(t.Lmp, _rows_of(P), out_exp="col", bt_rowexp=P.exp), -1, out_exp="col")
        P = add(P, gemm_nt(t.X, _rows_of(resid), out_exp="col", bt_rowexp=resid.exp), 1, out_exp="align")
    return P


def trsv_tile(t, Vlimbs, vexp, out_exp=None):
    """y = L^-1 v for many vectors with a factored tile (forward substitution on the GPU).
    Vlimbs: (L, nvec, b) limb planes, vexp: (nvec,) exponents, each vector fixed point with 32L - 40
    fractional bits of its own scale. The vectors are widened exactly to the tile's wide format
    (40 more fractional bits, entry-major words), solved, and rounded back: to headroom_bits(L)
    below each vector's largest entry (out_exp None) or to the int exponent out_exp.
    Returns (limbs (L, nvec, b), exps (nvec,))."""
    L, nvec, b = Vlimbs.shape
    widen = 32 * t.LF - (32 * L - 40)                          # extra fractional bits of the wide format
    WL = t.WL
    n = nvec * b
    dv = Vlimbs.device
    wide = shift_round(Vlimbs.reshape(L, n), torch.full((n,), -widen, dtype=torch.int64, device=dv), WL)
    entry_major = wide.view(WL, nvec, b).permute(1, 2, 0).contiguous()
    call("sdpv_tile_trsv", ptr(t.Lw), ptr(t.RINV), ptr(entry_major), nvec, b, WL, t.LF, stream())
    planes = entry_major.permute(2, 0, 1).reshape(WL, n)
    wide_exp = vexp - widen                                    # exponent of the wide words, per vector
    if out_exp is None:
        target = bitlen_words(planes.contiguous()).view(nvec, b).amax(1) + wide_exp - headroom_bits(L)
    else:
        target = torch.full((nvec,), int(out_exp), dtype=torch.int64, device=dv)
    shift = (target - wide_exp).repeat_interleave(b)
    return shift_round(planes.contiguous(), shift, L).view(L, nvec, b), target
