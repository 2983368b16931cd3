"""Solver-side helpers of the multiprecision core (mirror of parts of sdpvibe/solver.py): MPF block
helpers, the Cholesky condition estimate, panelled solves and the step length.

Step length: alpha = min(1, -gamma / lambda_min) with lambda_min the smallest eigenvalue over the
blocks of L^-1 dM L^-T. The eigenvalues come from a float64 copy of each block with the error bound
(n + 2) eps ||M||_F; blocks whose bound could move the step length by more than EIG_TOL are
recomputed by multiprecision deflation (refine_min_eig). The float64 copies, the symmetric
eigensolver and the float64 -> MPF conversion run on the GPU (gfx950 kernels); only per-block
scalars go to the host, where the step-length arithmetic is done in mpmath as in the reference.
"""
import os

import mpmath as mp
import torch

from . import mpf as F
from ._lib import call, ptr, stream
from .mpf import EZ, MPF, Chol, add, gemm, symmetrize


# ------------------------------------------------------------------ MPF block helpers (solver.py:39-79)
def put(dst, r0, c0, src):
    h, w = src.shape[-2:]
    dst.m[..., r0:r0 + h, c0:c0 + w] = src.m
    dst.e[..., r0:r0 + h, c0:c0 + w] = src.e


def sub(A, r0, r1, c0, c1):
    return MPF(A.m[..., r0:r1, c0:c1].contiguous(), A.e[..., r0:r1, c0:c1].contiguous())


def masked(A, M):
    """zero the entries where M (broadcastable bool) is False"""
    M = M.expand(A.shape)
    return MPF(torch.where(M[None], A.m, 0), torch.where(M, A.e, EZ))


def copy_lower_to_upper(A):
    n = A.shape[-1]
    iu = torch.triu_indices(n, n, 1, device=A.m.device)
    m, e = A.m.clone(), A.e.clone()
    m[..., iu[0], iu[1]] = A.m[..., iu[1], iu[0]]
    e[..., iu[0], iu[1]] = A.e[..., iu[1], iu[0]]
    return MPF(m, e)


def add_diag(A, s, dmask=None):
    """A + s on the diagonal (where dmask (nb, n) is True); s: MPF scalar"""
    n = A.shape[-1]
    idx = torch.arange(n, device=A.m.device)
    d = MPF(A.m[..., idx, idx].contiguous(), A.e[..., idx, idx].contiguous())
    sv = MPF(s.m.reshape(-1, *([1] * d.e.dim())).expand(d.m.shape).contiguous(),
             s.e.reshape([1] * d.e.dim()).expand(d.e.shape).contiguous())
    if dmask is not None:
        sv = masked(sv, dmask)
    d = add(d, sv)
    m, e = A.m.clone(), A.e.clone()
    m[..., idx, idx] = d.m
    e[..., idx, idx] = d.e
    return MPF(m, e)


# ------------------------------------------------------------------ condition estimate (solver.py:548-569)
def _diag_log2(ch):
    n, L = ch.n, ch.L
    idx = torch.arange(n, device=ch.Lt.limbs.device)
    d = ch.Lt.limbs[:, :, idx, idx]                                 # (L, nb, n) fixed point, > 0
    top = d[-1].double() * 2.0 ** 32 + (d[-2].double() % 2.0 ** 32) if L > 1 else d[-1].double()
    return torch.log2(torch.clamp(top, min=1e-300)) + (ch.Lt.exp + 32 * (L - 2)) + ch.k.double()


def chol_log2_cond_parts(ch):
    """(max, min) over the diagonal of L (log2) for a single-matrix Chol"""
    lg = _diag_log2(ch)
    return float(lg.max()), float(lg.min())


def chol_log2_cond(ch, mask=None):
    """per block log2 of the Cholesky condition estimate (max/min diagonal of L)^2; mask (nb, n):
    rows that count (padding excluded)"""
    lg = _diag_log2(ch)
    if mask is None:
        mask = torch.ones_like(lg, dtype=torch.bool)
    hi = torch.where(mask, lg, torch.full_like(lg, -1e300)).amax(1)
    lo = torch.where(mask, lg, torch.full_like(lg, 1e300)).amin(1)
    return torch.where(mask.any(1), 2 * (hi - lo), torch.zeros_like(hi))


def solve_panels(cs, B, panels):
    """L^-1 B in column panels (identical result: alignment and rounding are per column)
    (solver.py:985-993)"""
    N = B.shape[-1]
    if panels <= 1:
        return cs.solve_lower(B)
    w = -(-N // panels)
    P = F.zeros(B.shape, B.L)
    for j0 in range(0, N, w):
        put(P, 0, j0, cs.solve_lower(sub(B, 0, B.shape[-2], j0, min(N, j0 + w))))
    return P


# ------------------------------------------------------------------ float64 on the GPU
def to_double(A, shift=None):
    """MPF (nb, n, n) -> float64 tensor of A * 2^-shift[b] (MPF.to_float semantics)"""
    L = A.L
    m, e = A.m.contiguous(), A.e.contiguous()
    n = e.numel()
    out = torch.empty(A.shape, dtype=torch.float64, device=m.device)
    per = n // max(1, A.shape[0]) if shift is not None else 1
    sh = shift.contiguous() if shift is not None else None
    call("sdpv_mpf_to_double", ptr(m), ptr(e), L, n, ptr(sh), per, ptr(out), stream())
    return out


def symeig(Mf, vectors=False):
    """(w ascending (nb, n), V (nb, n, n) or None, ||(M + M^T)/2||_F (nb,)) of float64 (nb, n, n)"""
    nb, n, _ = Mf.shape
    m = n + (n & 1)
    Mf = Mf.contiguous()
    work = torch.empty((2 * nb * m * m,), dtype=torch.float64, device=Mf.device)
    w = torch.empty((nb, n), dtype=torch.float64, device=Mf.device)
    V = torch.empty((nb, n, n), dtype=torch.float64, device=Mf.device) if vectors else None
    fro = torch.empty((nb,), dtype=torch.float64, device=Mf.device)
    call("sdpv_symeig", ptr(Mf), nb, n, ptr(work), ptr(w), ptr(V), ptr(fro), stream())
    return w, V, fro


EIG_TOL = float(os.environ.get("SDPB_GPU_EIG_TOL", "1e-10"))       # relative accuracy of the step lengths
_EPS = 2.0 ** -52
_DEFLATE = 2.0 ** -20                                                # refine_min_eig: deflated eigenvalues > this ||A||_F


def _eig_err(fro, n):
    """bound on |computed - exact| of the float64 eigenvalues, per block: rounding to float64 plus the
    eigensolver's backward error, (n + 2) eps ||Mf||_F (solver.py:1002-1005)"""
    return (n + 2) * _EPS * fro


def _resolved(lam, err, gamma):
    """every eigenvalue in [lam - err, lam + err] gives the same step length to EIG_TOL"""
    return step_from_eig(lam - err, gamma) >= (1 - EIG_TOL) * step_from_eig(lam + err, gamma)


def _mpf_from_float(V, L):
    """float64 tensor (any shape, on the GPU) -> MPF, exactly (solver.py:1014-1020)"""
    v = V.contiguous().reshape(-1)
    n = v.numel()
    m = torch.empty((L, n), dtype=torch.int32, device=v.device)
    e = torch.empty((n,), dtype=torch.int64, device=v.device)
    call("sdpv_mpf_from_double", ptr(v), n, L, ptr(m), ptr(e), stream())
    return MPF(m.reshape(L, *V.shape), e.reshape(*V.shape))


def refine_min_eig(A, gamma, levels=64):
    """lambda_min of each symmetric MPF matrix A[b] (nb, n, n) to the accuracy the step length needs
    (EIG_TOL), for blocks whose float64 copy cannot resolve it (solver.py:1028-1061). Each level
    deflates the eigenvalues above _DEFLATE ||A||_F: with V the float64 eigenvectors,
    W = V^T A V (multiprecision) is congruent to A and nearly diagonal; the eigenvalues of W below the
    deflated ones are those of the Schur complement W_ss - W_sb W_bb^-1 W_bs, formed at full size
    (deflated rows and columns zero) so that the blocks stay batched. Returns (values, number of
    blocks left unresolved, whose value is the lower end of their float64 error interval)."""
    L = A.L
    out, idx, unresolved = [None] * A.shape[0], list(range(A.shape[0])), set()
    one = F.scalar(1, L)
    n = A.shape[-1]
    for _ in range(levels):
        s = F._top(A).reshape(len(idx), -1).amax(1)
        zero = (s <= EZ // 2).tolist()
        s = torch.where(s <= EZ // 2, torch.zeros_like(s), s)
        w, V, fro = symeig(to_double(A, s), vectors=True)
        err = _eig_err(fro, n)
        big = w > _DEFLATE * fro[:, None]
        wl, el, sl, bl = w[:, 0].tolist(), err.tolist(), s.tolist(), big.cpu().numpy()
        keep = []
        for j, q in enumerate(idx):
            sc = mp.mpf(2) ** int(sl[j])
            lam, e = mp.mpf(wl[j]) * sc, mp.mpf(el[j]) * sc
            unresolved.discard(q)
            if zero[j]:
                out[q] = mp.mpf(0)
            elif _resolved(lam, e, gamma):
                out[q] = lam
            else:
                out[q] = lam - e                                     # until resolved: the shorter, safe step
                unresolved.add(q)
                if bl[j].any() and not bl[j].all():
                    keep.append(j)
        if not keep:
            break
        kt = torch.tensor(keep, device=A.m.device)
        A, idx = A[kt], [idx[j] for j in keep]
        Pb = big[kt]
        Ps = ~Pb                                                     # deflated / kept rows
        Vm = _mpf_from_float(V[kt], L)
        W = symmetrize(gemm(Vm, gemm(A, Vm), transA=True))
        Wbb = add_diag(masked(W, Pb[:, :, None] & Pb[:, None, :]), one, Ps)   # identity on the kept rows
        Y = Chol(Wbb).solve_lower(masked(W, Pb[:, :, None] & Ps[:, None, :]))
        A = symmetrize(add(masked(W, Ps[:, :, None] & Ps[:, None, :]), gemm(Y, Y, transA=True), -1))
    return out, len(unresolved)


def min_eig(chols, dMs, lam=None, gamma=None, stats=None):
    """min over blocks of lambda_min(L^-1 dM L^-T), combined with lam (solver.py:1064-1093). With
    gamma (stepLengthReduction), blocks whose float64 error bound could move the step length by
    more than EIG_TOL are recomputed in multiprecision (counted in stats["eig_refined"])."""
    for ch, D in zip(chols, dMs):
        T = ch.solve_lower(D)
        M = ch.solve_lower(T.T)                                      # L^-1 dM L^-T (dM symmetric)
        tp = F._top(M)
        s = tp.reshape(tp.shape[0], -1).amax(1)                      # per block scale
        ok = s > EZ // 2
        s = torch.where(ok, s, torch.zeros_like(s))
        w, _, fro = symeig(to_double(M, s))
        n = M.shape[-1]
        ev, err, sl = w[:, 0].tolist(), _eig_err(fro, n).tolist(), s.tolist()
        flag = []
        for q in torch.nonzero(ok).reshape(-1).tolist():
            sc = mp.mpf(2) ** int(sl[q])
            lv = mp.mpf(ev[q]) * sc
            if gamma is not None and not _resolved(lv, mp.mpf(err[q]) * sc, gamma):
                flag.append(q)
                continue
            lam = lv if lam is None else min(lam, lv)
        if flag:
            vals, nu = refine_min_eig(M[torch.tensor(flag, device=M.m.device)], gamma)
            for lv in vals:
                lam = lv if lam is None else min(lam, lv)
            if stats is not None:
                stats["eig_refined"] = stats.get("eig_refined", 0) + len(flag)
                stats["eig_unresolved"] = stats.get("eig_unresolved", 0) + nu
    return lam


def step_from_eig(lam, gamma):
    """gamma * (-1/lambda_min) capped at 1"""
    if lam is None or lam > -gamma:
        return mp.mpf(1)
    return -gamma / lam


def step_length(chols, dMs, gamma):
    return step_from_eig(min_eig(chols, dMs), gamma)



def _rand_mpf(shape, L, gen, spread=4):
    """random normalised MPF entries of magnitude ~2^[-spread, spread] (solver.py:401-406): random
    limbs from the generator, normalised on the GPU"""
    n = 1
    for s in shape:
        n *= s
    m = torch.randint(-2 ** 31, 2 ** 31, (L, n), dtype=torch.int32, device=gen.device, generator=gen)
    e = torch.randint(-spread, spread + 1, (n,), dtype=torch.int64, device=gen.device, generator=gen) - 32 * L
    return F.normalize(m, e, L).reshape(*shape)
