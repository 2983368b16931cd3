"""Base case on the GPU: batched multiprecision Cholesky + triangular inverse of small tiles
(mirror of sdpvibe/basecase.py).

Wide fixed-point format inside the kernel: WL = L + 24 32-bit words, two's complement, F = 32L
fractional bits. Input tiles are SPD and equilibrated, so |a_ij| <= 1 and |l_ij| <= 1; pivots may
be tiny and their reciprocals (and X = L^-1) large, hence the integer words.
"""
import torch

from ._lib import call, ptr, stream
from .limbops import _errbuf

LI = 24   # integer words of the wide format


def tile_factor(limbs, BT):
    """limbs: (L, nb, BT, BT) fixed point at exponent -(32L-40) -> (Lout (L, nb, BT, BT) at the same
    exponent, upper triangle zero; Xw (L+24, nb, BT, BT) = L^-1 at exponent -32L)"""
    L, nb = limbs.shape[0], limbs.shape[1]
    WL = L + LI
    A = limbs.contiguous()
    Lout = torch.empty_like(A)
    Xw = torch.empty((WL, nb, BT, BT), dtype=torch.int32, device=A.device)
    work = torch.empty((3 * nb * BT * BT * WL,), dtype=torch.int32, device=A.device)
    call("sdpv_tile_chol", ptr(A), L, nb, BT, ptr(Lout), ptr(Xw), ptr(work), ptr(_errbuf()), stream())
    return Lout, Xw
