"""Sensitivities of an SDP's optimal value to its constraint data, from the solver's final iterate and
Cholesky factors (no counterpart in the reference, which drops its factors after every iteration,
sdpvibe/solver.py:946-947).

SDP (sdpvibe_amd.sdp): min c.x + const s.t. X = sum_p A_p x_p >= 0, B^T x = b (dual: max b.y + const,
Tr(A_p Y) + (B y)_p = c_p, Y >= 0). At a strictly complementary optimum the optimal value v has the
first-order expansion (the Lagrangian c.x - y.(B^T x - b) - Tr(Y (sum A_p x_p - X)) is stationary):

    dv = dc . x + db . y - x^T dB y            (1)

Adjoint of the Schur-complement system. The solver's Newton system at the iterate is
K [dx; dy] = [r1; r2], K = [[S, -B], [B^T, 0]], with S = L_S L_S^T per block and
Q = B^T S^-1 B = P^T P (P = L_S^-1 B). A data perturbation enters the right-hand side as
dr1 = -dc + dB y, dr2 = db - dB^T x, so for any linear functional l = g_x . x + g_y . y of the solution

    dl = lam_x . dr1 + lam_y . dr2,   K^T [lam_x; lam_y] = [g_x; g_y],
    lam_y = Q^-1 (g_y + P^T L_S^-1 g_x),   lam_x = L_S^-T (L_S^-1 g_x - P lam_y),

i.e. one solve with the stored factor of Q and two triangular solves per block: the gradients are
dl/dc = -lam_x, dl/db = lam_y, dl/dB = lam_x y^T - x lam_y^T. All products are exact multiprecision
GEMMs. With l = c.x (or b.y) this gives the gradient of the primal (dual) objective
(`objective_gradients`); `tangent` is the forward solve of the same system. `directional` evaluates
(1) for a batch of directions with one exact GEMM per term; `SDPValue` exposes
v(theta) = v(data + sum_k theta_k D_k) to torch.autograd, its backward through the adjoint solve.
"""
import mpmath as mp
import torch

from . import mpf as F
from .blockops import repeat_batch, stack_rows
from .mpf import MPF, add, gemm, mul, total


def _cat_rows(parts):
    return MPF(torch.cat([p.m for p in parts], 1), torch.cat([p.e for p in parts], 0))


def refactor(solver):
    """Cholesky factors of X, Y, S and Q at the solver's current iterate (the last computed step
    applied), kept in solver.final_factors together with P = L_S^-1 B per group"""
    solver._apply_update()
    r = solver._residues()
    cQ, _, _ = solver._schur_and_q(r["work"])
    w = r["work"]
    solver.final_factors = dict(X=[v["cX"] for v in w], Y=[v["cY"] for v in w], S=[v["cS"] for v in w],
                                P=[solver._P(g, v) for g, v in zip(solver.sdp.groups, w)], Q=cQ)
    return solver.final_factors


def primal_vector(solver):
    """x over all blocks, stacked in block order (sum P_j, 1)"""
    return _cat_rows([stack_rows(st.x) for st in solver.state])


def first_order(solver):
    """the gradients of (1) as MPF: (dv/dc = x (sum P, 1), dv/db = y (N, 1), dv/dB = -x y^T (sum P, N))"""
    x = primal_vector(solver)
    return x, solver.y, -gemm(x, solver.y, transB=True)


def directional(solver, dc, dB, db):
    """dv along K directions by (1): dc (K, sum P, 1), dB (K, sum P, N), db (K, N, 1) (MPF, rows in
    block order) -> list of K mpmath numbers; each is one `total` over the entrywise products
    dc x, db y and -x (dB y), the products exact GEMMs rounded once"""
    x = primal_vector(solver)
    K = dc.shape[0]
    xs, ys = repeat_batch(x, K), repeat_batch(solver.y, K)
    t1 = mul(dc, xs)
    t2 = mul(db, ys)
    t3 = -mul(gemm(dB, ys), xs)
    out = []
    for k in range(K):
        parts = [t1[k], t2[k], t3[k]]
        flat = MPF(torch.cat([p.m.reshape(p.L, -1) for p in parts], 1), torch.cat([p.e.reshape(-1) for p in parts]))
        out.append(total(flat))
    return out


def _group_slices(solver, v):
    """a vector over all blocks (sum P, C) -> per group (nb, P, C) MPF"""
    out, o = [], 0
    for g in solver.sdp.groups:
        n = g.nb * g.P
        out.append(MPF(v.m[:, o:o + n].contiguous(), v.e[o:o + n].contiguous()).reshape(g.nb, g.P, v.shape[-1]))
        o += n
    return out


def _factors(solver):
    ff = solver.final_factors
    return ff if ff is not None and "P" in ff else refactor(solver)


def adjoint(solver, gx, gy):
    """gradients of l = gx . x + gy . y (gx (sum P, 1), gy (N, 1) MPF) with respect to (c, b, B) at the
    factored iterate, through the adjoint Schur-complement solve (see the module docstring)"""
    ff = _factors(solver)
    groups = solver.sdp.groups
    u = [cS.solve_lower(v) for cS, v in zip(ff["S"], _group_slices(solver, gx))]     # L_S^-1 g_x
    Pall = _cat_rows([stack_rows(P) for P in ff["P"]])
    uall = _cat_rows([stack_rows(v) for v in u])
    lam_y = ff["Q"].solve(add(gy, gemm(Pall, uall, transA=True)))
    lam_x = []
    for g, cS, P, v in zip(groups, ff["S"], ff["P"], u):
        lam_x.append(stack_rows(cS.solve_upper(add(v, gemm(P, repeat_batch(lam_y, g.nb)), -1))))
    lam_x = _cat_rows(lam_x)
    x = primal_vector(solver)
    dB = add(gemm(lam_x, solver.y, transB=True), gemm(x, lam_y, transB=True), -1)
    return -lam_x, lam_y, dB


def tangent(solver, dc, dB, db):
    """(dx, dy) solving the Newton system K [dx; dy] = [-dc + dB y; db - dB^T x] with the stored factors
    (dc (sum P, 1), dB (sum P, N), db (N, 1)): dy = Q^-1 (r2 - P^T L_S^-1 r1), dx = L_S^-T (L_S^-1 r1 + P dy).
    adjoint() is its transpose: gx . dx + gy . dy equals the contraction of adjoint(gx, gy) with the
    direction."""
    ff = _factors(solver)
    x = primal_vector(solver)
    r1 = add(gemm(dB, solver.y), dc, -1)
    r2 = add(db, gemm(dB, x, transA=True), -1)
    u = [cS.solve_lower(v) for cS, v in zip(ff["S"], _group_slices(solver, r1))]
    Pall = _cat_rows([stack_rows(P) for P in ff["P"]])
    dy = ff["Q"].solve(add(r2, gemm(Pall, _cat_rows([stack_rows(v) for v in u]), transA=True), -1))
    dx = [stack_rows(cS.solve_upper(add(v, gemm(P, repeat_batch(dy, g.nb)))))
          for g, cS, P, v in zip(solver.sdp.groups, ff["S"], ff["P"], u)]
    return _cat_rows(dx), dy


def c_vector(sdp):
    """c over all blocks (sum P, 1), in the order of primal_vector"""
    L = sdp.L
    return MPF(torch.cat([g.c.m.reshape(L, -1) for g in sdp.groups], 1).reshape(L, -1, 1),
               torch.cat([g.c.e.reshape(-1) for g in sdp.groups]).reshape(-1, 1))


def objective_gradients(solver, which="primal"):
    """(dv/dc (sum P, 1), dv/db (N, 1), dv/dB (sum P, N)) of the primal objective c.x (which="primal") or
    the dual objective b.y ("dual") at the solver's iterate, through the adjoint solve:
    d(c.x) = dc.x + c.dx, d(b.y) = db.y + b.dy, with c.dx and b.dy from adjoint()"""
    x, y = primal_vector(solver), solver.y
    L = solver.L
    if which == "primal":
        gc, gb, gB = adjoint(solver, c_vector(solver.sdp), F.zeros(tuple(y.shape), L))
        return add(gc, x), gb, gB
    if which == "dual":
        gc, gb, gB = adjoint(solver, F.zeros(tuple(x.shape), L), solver.sdp.b)
        return gc, add(gb, y), gB
    raise ValueError(f"which must be 'primal' or 'dual', not {which!r}")


def contract(grads, dc, dB, db):
    """sum of grads . direction for each of K directions (dc (K, sum P, 1), dB (K, sum P, N),
    db (K, N, 1)) -> list of K mpmath numbers, each product exact and the sum one `total`"""
    gc, gb, gB = grads
    K = dc.shape[0]
    t1 = mul(dc, repeat_batch(gc, K))
    t2 = mul(db, repeat_batch(gb, K))
    t3 = mul(dB, repeat_batch(gB, K))
    out = []
    for k in range(K):
        parts = [t1[k], t2[k], t3[k]]
        out.append(total(MPF(torch.cat([q.m.reshape(q.L, -1) for q in parts], 1),
                             torch.cat([q.e.reshape(-1) for q in parts]))))
    return out


class SDPValue(torch.autograd.Function):
    """v(theta) = optimal value of the SDP with data (c, B, b) + sum_k theta_k (dc_k, dB_k, db_k), for a
    solver already run to optimality at theta = 0. forward returns v (the primal or the dual objective)
    as a float64 scalar (its exact mpmath value is kept in ctx.value); backward returns dv/dtheta_k,
    by default through the adjoint Schur-complement solve (objective_gradients), or from the
    first-order estimate (1) with method="first_order".
    Use: v = SDPValue.apply(theta, solver, (dc, dB, db), "primal", "adjoint")."""

    @staticmethod
    def forward(ctx, theta, solver, directions, which="primal", method="adjoint"):
        if method not in ("adjoint", "first_order"):
            raise ValueError(f"method must be 'adjoint' or 'first_order', not {method!r}")
        ctx.solver, ctx.directions, ctx.which, ctx.method = solver, directions, which, method
        ctx.value = solver.result["primalObjective" if which == "primal" else "dualObjective"]
        return torch.tensor(float(ctx.value), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        if ctx.method == "adjoint":
            d = contract(objective_gradients(ctx.solver, ctx.which), *ctx.directions)
        else:
            d = directional(ctx.solver, *ctx.directions)
        g = torch.tensor([float(v) for v in d], dtype=torch.float64) * grad
        return g, None, None, None, None


def perturbed(sdp, dc, dB, db, t):
    """a copy of the SDP with data (c, B, b) + t (dc, dB, db): dc (sum P, 1), dB (sum P, N), db (N, 1) MPF
    with rows in block order, t an mpmath number; each perturbed entry is rounded once (mpf add of
    the exact product t d rounded once)"""
    import copy
    L = sdp.L
    ts = F.scalar(t, L)
    new = copy.copy(sdp)
    new.groups = []
    new.store = None                                  # the perturbed copy keeps everything on the device
    o = 0
    for g in sdp.groups:
        n = g.nb * g.P
        ng = copy.copy(g)
        sl = lambda X: MPF(X.m[:, o:o + n].contiguous(), X.e[o:o + n].contiguous())
        ng.c = add(g.c, mul(sl(dc), ts).reshape(g.nb, g.P, 1))
        ng.B = add(sdp.B_of(g), mul(sl(dB), ts).reshape(g.nb, g.P, sdp.N))
        new.groups.append(ng)
        o += n
    new.b = add(sdp.b, mul(db, ts))
    return new
