"""Limb-level bookkeeping in limb-plane layout (mirror of sdpvibe/limbops.py).

A batch of B multiword integers is an int32 tensor X of shape (L, B): X[k, e] is the k-th 32-bit
limb (little endian, two's complement) of entry e. Every function launches one gfx950 kernel of
libsdpvibe_hip.so on the current stream; none synchronises with the host except
shift_round(..., check=True), which must inspect its flags.
"""
import torch

from ._lib import call, ptr, stream

dev = "cuda"

# device -> sticky overflow flag (int32 (1,)), set by rounding kernels (sdpvibe/limbops.py:156-169)
ERR = {}


def _errbuf():
    d = torch.cuda.current_device()
    if d not in ERR:
        ERR[d] = torch.zeros((1,), dtype=torch.int32, device=dev)
    return ERR[d]


def check_errors():
    """Raise OverflowError if a rounding step overflowed since the last check (limbops.py:165-169)."""
    e = ERR.get(torch.cuda.current_device())
    if e is not None and int(e[0]):
        e.zero_()
        raise OverflowError("a rounding step overflowed its limb budget")


def shift_round(X, s, Lout, check=False, out=None):
    """X: (Lin, B) int32 limbs; s: (B,) int64 -> (Lout, B) = round_half_up(X / 2^s)
    (limbops.py:173-185); into `out` when given (contiguous (Lout, B) int32, not overlapping X)."""
    Lin, B = X.shape
    X = X.contiguous()
    s = s.contiguous()
    if out is None:
        out = torch.empty((Lout, B), dtype=torch.int32, device=X.device)
    elif tuple(out.shape) != (Lout, B) or out.dtype != torch.int32 or not out.is_contiguous():
        raise ValueError(f"shift_round: out must be contiguous int32 {(Lout, B)}")
    flag = torch.zeros((B,), dtype=torch.int32, device=X.device) if check else _errbuf()
    call("sdpv_shift_round", ptr(X), Lin, B, ptr(s), ptr(out), Lout, ptr(flag), int(check), stream())
    if check and bool(flag.any()):
        raise OverflowError("result does not fit in the requested number of limbs")
    return out


def add_same(A, Bm, sgn=1):
    """A + sgn*B for (L, B) int32 limbs with identical exponents (wraps mod 2^(32L))
    (limbops.py:188-193)."""
    L, B = A.shape
    A, Bm = A.contiguous(), Bm.contiguous()
    out = torch.empty_like(A)
    call("sdpv_add_same", ptr(A), ptr(Bm), ptr(out), L, B, int(sgn), stream())
    return out


def bitlen_words(X):
    """(L, B) int32 limbs -> (B,) int64 bit length of v (of ~v when v < 0) (limbops.py:213-218)."""
    L, B = X.shape
    X = X.contiguous()
    out = torch.empty((B,), dtype=torch.int64, device=X.device)
    call("sdpv_bitlen", ptr(X), L, B, ptr(out), stream())
    return out
