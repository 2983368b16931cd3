"""Block-structured views and edits of MPF matrices, used by the interior-point driver and the step
length. Every helper returns a new MPF (or writes a window in place, `paste`); entries are moved,
never rounded, except where a helper says it adds or multiplies.

Batched matrices are MPF of shape (nb, R, C): limbs (L, nb, R, C), exponents (nb, R, C). Index
tensors are built on the device of the operand, so nothing here synchronises with the host.
"""
import torch

from .mpf import EZ, MPF, add, mul


def window(A, r0, r1, c0, c1):
    """copy of rows r0:r1, columns c0:c1 of the last two dimensions"""
    return MPF(A.m[..., r0:r1, c0:c1].contiguous(), A.e[..., r0:r1, c0:c1].contiguous())


def paste(dst, r0, c0, src):
    """write src into dst at rows r0.., columns c0.. (in place)"""
    h, w = src.shape[-2:]
    dst.m[..., r0:r0 + h, c0:c0 + w] = src.m
    dst.e[..., r0:r0 + h, c0:c0 + w] = src.e


def keep(A, where):
    """A with the entries where `where` (broadcastable bool) is False set to zero"""
    w = where.expand(A.shape)
    return MPF(A.m * w.to(A.m.dtype)[None], torch.where(w, A.e, torch.full_like(A.e, EZ)))


def mirror_lower(A):
    """symmetric matrix from the lower triangle of A (the strict upper triangle is overwritten)"""
    n = A.shape[-1]
    r = torch.arange(n, device=A.e.device)
    up = r[:, None] < r[None, :]
    return MPF(torch.where(up, A.m.transpose(-1, -2), A.m), torch.where(up, A.e.transpose(-1, -2), A.e))


def diagonal(A):
    """(..., n, n) -> (..., n, 1): the diagonal as a column"""
    m = torch.diagonal(A.m, dim1=-2, dim2=-1)
    e = torch.diagonal(A.e, dim1=-2, dim2=-1)
    return MPF(m.unsqueeze(-1).contiguous(), e.unsqueeze(-1).contiguous())


def _with_diagonal(A, d):
    """A with its diagonal replaced by the column d (..., n, 1)"""
    n = A.shape[-1]
    eye = torch.eye(n, dtype=torch.bool, device=A.e.device)
    return MPF(torch.where(eye, d.m.transpose(-1, -2), A.m), torch.where(eye, d.e.transpose(-1, -2), A.e))


def _scalar_column(s, like):
    """the MPF scalar s repeated in the shape of the column `like`"""
    L = s.L
    m = s.m.reshape(L, *([1] * like.e.dim())).expand(like.m.shape).contiguous()
    e = s.e.reshape([1] * like.e.dim()).expand(like.e.shape).contiguous()
    return MPF(m, e)


def shift_diagonal(A, s, rows=None):
    """A + s I restricted to the diagonal entries whose row is selected by rows (nb, n) (all when
    None); s an MPF scalar; each diagonal entry rounded once (mpf add)"""
    d = diagonal(A)
    sv = _scalar_column(s, d)
    if rows is not None:
        sv = keep(sv, rows.unsqueeze(-1))
    return _with_diagonal(A, add(d, sv))


def fill_diagonal(A, s, rows):
    """A with the diagonal entries of the rows selected by rows (nb, n) set to the MPF scalar s"""
    d = diagonal(A)
    sv = _scalar_column(s, d)
    sel = rows.unsqueeze(-1)
    return _with_diagonal(A, MPF(torch.where(sel[None], sv.m, d.m), torch.where(sel, sv.e, d.e)))


def scale_columns(V, v):
    """V (nb, h, K) with column k multiplied by v[b, k] (v: (nb, K, 1)); each product rounded once"""
    nb, h, K = V.shape
    w = MPF(v.m.reshape(v.L, nb, 1, K).expand(-1, -1, h, -1).contiguous(),
            v.e.reshape(nb, 1, K).expand(-1, h, -1).contiguous())
    return mul(V, w)


def repeat_batch(v, nb):
    """(R, C) -> (nb, R, C): the same matrix for every block"""
    return MPF(v.m.unsqueeze(1).expand(-1, nb, -1, -1).contiguous(), v.e.unsqueeze(0).expand(nb, -1, -1).contiguous())


def twice(a):
    """(nb, R, C) -> (2 nb, R, C): the same vector for the even and the odd parity"""
    return MPF(a.m.repeat(1, 2, 1, 1), a.e.repeat(2, 1, 1))


def parity_sum(A, nb):
    """(2 nb, ...) -> (nb, ...): the even half plus the odd half (mpf add)"""
    return add(A[:nb], A[nb:])


def stack_rows(A):
    """(nb, R, C) -> (nb R, C)"""
    nb, R, C = A.shape
    return A.reshape(nb * R, C)
