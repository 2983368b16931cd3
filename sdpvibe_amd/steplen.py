"""Step length of the interior-point iteration: alpha = min(1, -gamma / lambda_min), lambda_min the
smallest eigenvalue over the blocks of L^-1 dM L^-T (X = L L^T), with the acceptance rule of the
reference (sdpvibe/solver.py:996-1103): a block's float64 eigenvalue is used when every value in its
error interval gives the same step length to EIG_TOL; otherwise the block is refined in
multiprecision by deflating its large eigenvalues.

Design. Each block is scaled by a power of two (its largest entry), copied to float64 on the GPU and
diagonalised by the batched Jacobi kernel, which also reports the Frobenius norm and what it left
off the diagonal; the host receives three numbers per block and keeps an interval
[lam - err, lam + err] per block in mpmath. Refinement is a loop over levels that shrinks the batch
to the blocks still unresolved.
"""
import os

import mpmath as mp
import torch

from . import mpf as F
from ._lib import call, ptr, stream
from .blockops import keep, shift_diagonal
from .mpf import EZ, MPF, Chol, add, gemm, symmetrize

EIG_TOL = float(os.environ.get("SDPB_GPU_EIG_TOL", "1e-10"))       # relative accuracy of the step lengths
_EPS = 2.0 ** -52
_DEFLATE = 2.0 ** -20              # refinement: eigenvalues above this fraction of ||A||_F are deflated


# ------------------------------------------------------------------ float64 on the GPU
def to_double(A, shift=None):
    """MPF (nb, ...) -> float64 tensor of A * 2^-shift[b] (MPF.to_float's rounding)"""
    m, e = A.m.contiguous(), A.e.contiguous()
    n = e.numel()
    out = torch.empty(A.shape, dtype=torch.float64, device=m.device)
    per = n // max(1, A.shape[0]) if shift is not None else 1
    sh = shift.contiguous() if shift is not None else None
    call("sdpv_mpf_to_double", ptr(m), ptr(e), A.L, n, ptr(sh), per, ptr(out), stream())
    return out


def _mpf_from_float(V, L):
    """float64 tensor on the GPU -> MPF, exactly (53-bit mantissas)"""
    v = V.contiguous().reshape(-1)
    n = v.numel()
    m = torch.empty((L, n), dtype=torch.int32, device=v.device)
    e = torch.empty((n,), dtype=torch.int64, device=v.device)
    call("sdpv_mpf_from_double", ptr(v), n, L, ptr(m), ptr(e), stream())
    return MPF(m.reshape(L, *V.shape), e.reshape(*V.shape))


def symeig(Mf, vectors=False):
    """eigen-decomposition of the symmetrised float64 (nb, n, n): (w ascending (nb, n), V (nb, n, n) or
    None, ||(M + M^T)/2||_F (nb,), off-diagonal norm left by the iteration (nb,))"""
    nb, n, _ = Mf.shape
    m = n + (n & 1)
    Mf = Mf.contiguous()
    dev = Mf.device
    work = torch.empty((2 * nb * m * m,), dtype=torch.float64, device=dev)
    w = torch.empty((nb, n), dtype=torch.float64, device=dev)
    V = torch.empty((nb, n, n), dtype=torch.float64, device=dev) if vectors else None
    fro = torch.empty((nb,), dtype=torch.float64, device=dev)
    off = torch.empty((nb,), dtype=torch.float64, device=dev)
    call("sdpv_symeig", ptr(Mf), nb, n, ptr(work), ptr(w), ptr(V), ptr(fro), ptr(off), stream())
    return w, V, fro, off


# ------------------------------------------------------------------ acceptance rule
def step_from_eig(lam, gamma):
    """gamma * (-1 / lambda_min), capped at 1 (also for no eigenvalue at all)"""
    if lam is None or lam > -gamma:
        return mp.mpf(1)
    return -gamma / lam


def _resolved(lam, err, gamma):
    """True when every eigenvalue in [lam - err, lam + err] gives the same step length to EIG_TOL"""
    return step_from_eig(lam - err, gamma) >= (1 - EIG_TOL) * step_from_eig(lam + err, gamma)


class _Float64Spectra:
    """float64 spectra of a batch of symmetric MPF matrices A (nb, n, n), each scaled by 2^-s with s
    its largest entry's binary exponent: per block the smallest eigenvalue and its error bound in
    units of 2^s, (n + 2) eps ||A_s||_F (rounding to float64 plus the eigensolver's backward error)
    plus the off-diagonal norm the Jacobi iteration left."""

    def __init__(self, A, vectors=False):
        nb, n = A.shape[0], A.shape[-1]
        top = F._top(A).reshape(nb, -1).amax(1)
        self.zero = top <= EZ // 2                                   # all-zero blocks
        self.scale = torch.where(self.zero, torch.zeros_like(top), top)
        w, self.V, fro, off = symeig(to_double(A, self.scale), vectors)
        self.w, self.fro = w, fro
        self.err = (n + 2) * _EPS * fro + off

    def intervals(self):
        """per block (lam, err) as mpmath numbers, or None for an all-zero block"""
        lo, er, sc, zr = (self.w[:, 0].tolist(), self.err.tolist(), self.scale.tolist(), self.zero.tolist())
        out = []
        for j in range(len(lo)):
            if zr[j]:
                out.append(None)
                continue
            p = mp.mpf(2) ** int(sc[j])
            out.append((mp.mpf(lo[j]) * p, mp.mpf(er[j]) * p))
        return out


# ------------------------------------------------------------------ multiprecision refinement
def _deflate(A, V, big, L):
    """one deflation level on the batch: W = V^T A V in multiprecision (congruent to A, nearly
    diagonal), then the Schur complement of the rows `big` (nb, n): W_ss - W_sb W_bb^-1 W_bs, formed
    at full size (deflated rows and columns zero) so the blocks stay batched"""
    small = ~big
    Vm = _mpf_from_float(V, L)
    W = symmetrize(gemm(Vm, gemm(A, Vm), transA=True))
    Wbb = shift_diagonal(keep(W, big[:, :, None] & big[:, None, :]), F.scalar(1, L), small)
    T = Chol(Wbb).solve_lower(keep(W, big[:, :, None] & small[:, None, :]))
    return symmetrize(add(keep(W, small[:, :, None] & small[:, None, :]), gemm(T, T, transA=True), -1))


def refine_min_eig(A, gamma, levels=64):
    """lambda_min of each symmetric MPF block A[b] (nb, n, n) to the accuracy the step length needs,
    for blocks whose float64 copy leaves it unresolved. Each level diagonalises the active blocks in
    float64; a block is done when its interval is resolved (or it is zero); otherwise its eigenvalues
    above _DEFLATE ||A||_F are deflated exactly (congruence by the float64 eigenvectors, then a Schur
    complement, whose extra zero eigenvalues cannot shorten the step) and it stays active. Returns
    (values, number of blocks still unresolved, which keep the lower end of their interval: the
    shorter, safe step)."""
    L = A.L
    nb = A.shape[0]
    result = [None] * nb
    open_ = set()
    active = list(range(nb))
    for _ in range(levels):
        sp = _Float64Spectra(A, vectors=True)
        big = sp.w > _DEFLATE * sp.fro[:, None]
        nbig = big.sum(1).tolist()
        n = A.shape[-1]
        survivors = []
        for j, (b, iv) in enumerate(zip(active, sp.intervals())):
            open_.discard(b)
            if iv is None:
                result[b] = mp.mpf(0)
                continue
            lam, err = iv
            if _resolved(lam, err, gamma):
                result[b] = lam
                continue
            result[b] = lam - err
            open_.add(b)
            if 0 < nbig[j] < n:                                       # something to deflate, something left
                survivors.append(j)
        if not survivors:
            break
        sel = torch.tensor(survivors, device=A.m.device)
        A = _deflate(A[sel], sp.V[sel], big[sel], L)
        active = [active[j] for j in survivors]
    return result, len(open_)


def min_eig(chols, dMs, lam=None, gamma=None, stats=None):
    """min over all blocks of lambda_min(L^-1 dM L^-T), combined with lam. chols[i] factors the batch
    X_i = L L^T, dMs[i] the matching symmetric directions. With gamma (stepLengthReduction) the
    blocks whose float64 interval is unresolved are refined (counted in stats["eig_refined"] and
    stats["eig_unresolved"])."""
    def lower(a, b):
        return b if a is None else min(a, b)

    for ch, dM in zip(chols, dMs):
        M = ch.solve_lower(ch.solve_lower(dM).T)                      # L^-1 dM L^-T (dM symmetric)
        sp = _Float64Spectra(M)
        pending = []
        for j, iv in enumerate(sp.intervals()):
            if iv is None:
                continue
            if gamma is not None and not _resolved(iv[0], iv[1], gamma):
                pending.append(j)
            else:
                lam = lower(lam, iv[0])
        if pending:
            vals, nu = refine_min_eig(M[torch.tensor(pending, device=M.m.device)], gamma)
            for v in vals:
                lam = lower(lam, v)
            if stats is not None:
                stats["eig_refined"] = stats.get("eig_refined", 0) + len(pending)
                stats["eig_unresolved"] = stats.get("eig_unresolved", 0) + nu
    return lam


def step_length(chols, dMs, gamma):
    return step_from_eig(min_eig(chols, dMs), gamma)
