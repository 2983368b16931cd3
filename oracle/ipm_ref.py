"""TEST INFRASTRUCTURE ONLY: SDPB's primal-dual interior-point iteration (Mehrotra predictor-
corrector), restated densely in mpmath as an independent CPU path for the GPU driver
(sdpvibe_amd/ipm.py). Imported by tests/ and tests/golden/ scripts only.

It follows the reference's statement of the algorithm (sdpvibe/solver.py:680-982; SDPB 3.x
src/sdp_solve): residues (:721-764), termination tests (:798-816), predictor and corrector
directions (:871-933), centering parameters (:935-940) and step lengths (:944-951, min over blocks of
lambda_min(L^-1 dM L^-T) scaled by stepLengthReduction, capped at 1). It shares no formulation with the
GPU path: every constraint matrix A_p is an explicit dense matrix, the Schur complement is
S_pq = sum_parity Tr(A_p X^-1 A_q Y) entry by entry, the Newton system [[S, -B], [B^T, 0]] is solved
by one LU factorisation, and eigenvalues come from mpmath's symmetric eigensolver. Parities with
no basis rows are left out instead of padded.
"""
import mpmath as mp

DEFAULTS = dict(maxIterations=500, dualityGapThreshold="1e-30", primalErrorThreshold="1e-30",
                dualErrorThreshold="1e-30", initialMatrixScalePrimal="1e20", initialMatrixScaleDual="1e20",
                feasibleCenteringParameter="0.1", infeasibleCenteringParameter="0.3", stepLengthReduction="0.7",
                maxComplementarity="1e100")


def segments(d):
    return [(c, r) for c in range(d) for r in range(c + 1)]


class Block:
    """one block: dim d, K points, bases Ve (he x K), Vo (ho x K); the explicit A_p per parity"""

    def __init__(self, dim, K, Ve, Vo, c, B, prec=1600):
        with mp.workprec(prec):
            self._build(dim, K, Ve, Vo, c, B)

    def _build(self, dim, K, Ve, Vo, c, B):
        self.d, self.K = dim, K
        self.c = [mp.mpf(v) for v in c]
        self.B = [[mp.mpf(v) for v in row] for row in B]
        self.bases = [[[mp.mpf(v) for v in row] for row in V] for V in (Ve, Vo) if V is not None and len(V) > 0]
        self.P = K * dim * (dim + 1) // 2
        self.A = []                                     # A[par][p]: dense (d h) x (d h)
        for V in self.bases:
            h = len(V)
            mats = []
            for (cc, r) in segments(dim):
                for k in range(K):
                    M = mp.zeros(dim * h, dim * h)
                    f = mp.mpf(1) if cc == r else mp.mpf(1) / 2
                    for i in range(h):
                        for j in range(h):
                            v = f * V[i][k] * V[j][k]
                            M[r * h + i, cc * h + j] += v
                            if cc != r:
                                M[cc * h + i, r * h + j] += v
                    mats.append(M)
            self.A.append(mats)

    def sizes(self):
        return [self.d * len(V) for V in self.bases]


def _tr(A, B):
    """Tr(A B)"""
    n = A.rows
    return mp.fsum(A[i, j] * B[j, i] for i in range(n) for j in range(n))


def _inner(A, B):
    n = A.rows
    return mp.fsum(A[i, j] * B[i, j] for i in range(n) for j in range(n))


def _sym(A):
    return (A + A.T) / 2


def _maxabs(vals):
    return max([abs(v) for v in vals] + [mp.mpf(0)])


def _lam_min(X, dX):
    """smallest eigenvalue of L^-1 dX L^-T, X = L L^T"""
    Lc = mp.cholesky(X)
    Li = mp.inverse(Lc)
    M = _sym(Li * dX * Li.T)
    return min(mp.eigsy(M, eigvals_only=True))


def solve(blocks, b, const=0, max_iter=20, prec=1600, **params):
    """Run the iteration on the SDP (blocks: list of Block; b: list of N numbers). Returns the list of
    per-iteration records (mpmath values) and the result record."""
    old = mp.mp.prec
    mp.mp.prec = prec
    try:
        return _solve(blocks, [mp.mpf(v) for v in b], mp.mpf(const), max_iter, params)
    finally:
        mp.mp.prec = old


def _solve(blocks, b, const, max_iter, params):
    p = dict(DEFAULTS)
    p.update(params)
    p = {k: (mp.mpf(v) if isinstance(v, str) else v) for k, v in p.items()}
    N = len(b)
    rows_total = sum(sum(bl.sizes()) for bl in blocks)
    x = [[mp.mpf(0)] * bl.P for bl in blocks]
    X = [[mp.eye(n) * p["initialMatrixScalePrimal"] for n in bl.sizes()] for bl in blocks]
    Y = [[mp.eye(n) * p["initialMatrixScaleDual"] for n in bl.sizes()] for bl in blocks]
    y = [mp.mpf(0)] * N
    alphaP = alphaD = mp.mpf(0)
    log = []
    it = 0
    while True:
        it += 1
        # ---- residues
        dobj = const + mp.fsum(bi * yi for bi, yi in zip(b, y))
        pobj = const + mp.fsum(cp * xp for bl, xb in zip(blocks, x) for cp, xp in zip(bl.c, xb))
        dres, Pres, mXY = [], [], []
        for bl, xb, Xb, Yb in zip(blocks, x, X, Y):
            dr = []
            for q in range(bl.P):
                t = mp.fsum(_tr(bl.A[par][q], Yb[par]) for par in range(len(bl.bases)))
                by = mp.fsum(bl.B[q][n] * y[n] for n in range(N))
                dr.append(bl.c[q] - t - by)
            dres.append(dr)
            Pres.append([sum((bl.A[par][q] * xb[q] for q in range(bl.P)), mp.zeros(Xb[par].rows, Xb[par].rows))
                         - Xb[par] for par in range(len(bl.bases))])
            mXY.append([-(Xb[par] * Yb[par]) for par in range(len(bl.bases))])
        pvec = [b[n] - mp.fsum(bl.B[q][n] * xb[q] for bl, xb in zip(blocks, x) for q in range(bl.P)) for n in range(N)]
        dual_err = _maxabs([v for dr in dres for v in dr])
        primal_err_P = _maxabs([M[i, j] for Pb in Pres for M in Pb for i in range(M.rows) for j in range(M.cols)])
        primal_err_p = _maxabs(pvec)
        primal_err = max(primal_err_P, primal_err_p)
        gap = abs(pobj - dobj) / max(abs(pobj) + abs(dobj), mp.mpf(1))
        dual_feas = dual_err < p["dualErrorThreshold"]
        primal_feas = primal_err < p["primalErrorThreshold"]
        pd_feas = dual_feas and primal_feas
        reason = None
        if pd_feas and gap < p["dualityGapThreshold"]:
            reason = "found primal-dual optimal solution"
        elif it > max_iter:
            reason = "maxIterations exceeded"
        if reason is None:
            mu = mp.fsum(_tr(Xb[par], Yb[par]) for Xb, Yb in zip(X, Y) for par in range(len(Xb))) / rows_total
            if mu > p["maxComplementarity"]:
                reason = "maxComplementarity exceeded"
        if reason:
            return log, dict(terminateReason=reason, primalObjective=pobj, dualObjective=dobj, dualityGap=gap,
                             primalError=primal_err, dualError=dual_err, iterations=it - 1, x=x, y=y, X=X, Y=Y)
        # ---- Schur complement (per block) and the Newton system's LU factors
        Xinv = [[mp.inverse(Xb[par]) for par in range(len(Xb))] for Xb in X]
        Ptot = sum(bl.P for bl in blocks)
        KKT = mp.zeros(Ptot + N, Ptot + N)
        off = 0
        for bl, Wb, Yb in zip(blocks, Xinv, Y):
            for par in range(len(bl.bases)):
                Mq = [Wb[par] * bl.A[par][q] * Yb[par] for q in range(bl.P)]
                for i in range(bl.P):
                    for j in range(bl.P):
                        KKT[off + i, off + j] += _tr(bl.A[par][i], Mq[j])
            for i in range(bl.P):
                for n in range(N):
                    KKT[off + i, Ptot + n] = -bl.B[i][n]
                    KKT[Ptot + n, off + i] = bl.B[i][n]
            off += bl.P

        def direction(beta, dpred):
            """solve for (dx, dy, dX, dY) with R = -XY (- dXp dYp) + beta mu I"""
            rhs, Rs = [], []
            for bi, bl in enumerate(blocks):
                Rb = []
                for par in range(len(bl.bases)):
                    R = mXY[bi][par] + beta * mu * mp.eye(X[bi][par].rows)
                    if dpred is not None:
                        R = R - dpred[0][bi][par] * dpred[1][bi][par]
                    Rb.append(R)
                Rs.append(Rb)
                Z = [_sym(Xinv[bi][par] * (Pres[bi][par] * Y[bi][par] - Rb[par])) for par in range(len(bl.bases))]
                for q in range(bl.P):
                    rhs.append(-dres[bi][q] - mp.fsum(_tr(bl.A[par][q], Z[par]) for par in range(len(bl.bases))))
            sol = mp.lu_solve(KKT, mp.matrix(rhs + pvec))
            dx, o = [], 0
            for bl in blocks:
                dx.append([sol[o + q] for q in range(bl.P)])
                o += bl.P
            dy = [sol[Ptot + n] for n in range(N)]
            dX, dY = [], []
            for bi, bl in enumerate(blocks):
                dXb, dYb = [], []
                for par in range(len(bl.bases)):
                    D = sum((bl.A[par][q] * dx[bi][q] for q in range(bl.P)), mp.zeros(X[bi][par].rows, X[bi][par].rows))
                    D = D + Pres[bi][par]
                    dXb.append(D)
                    dYb.append(_sym(Xinv[bi][par] * (Rs[bi][par] - D * Y[bi][par])))
                dX.append(dXb)
                dY.append(dYb)
            return dx, dy, dX, dY, Rs

        R_err = _maxabs([(M + mu * mp.eye(M.rows))[i, j] for Mb in mXY for M in Mb
                         for i in range(M.rows) for j in range(M.cols)])
        beta_p = mp.mpf(0) if pd_feas else p["infeasibleCenteringParameter"]
        dxp, dyp, dXp, dYp, _ = direction(beta_p, None)
        fr = mp.fsum(_inner(X[bi][par] + dXp[bi][par], Y[bi][par] + dYp[bi][par])
                     for bi in range(len(blocks)) for par in range(len(X[bi])))
        r = fr / (mu * rows_total)
        beta = r * r if r < 1 else r
        beta_c = min(max(p["feasibleCenteringParameter"], beta), mp.mpf(1)) if pd_feas \
            else max(p["infeasibleCenteringParameter"], beta)
        dx, dy, dX, dY, _ = direction(beta_c, (dXp, dYp))
        gm = p["stepLengthReduction"]
        lamP = min(_lam_min(X[bi][par], dX[bi][par]) for bi in range(len(blocks)) for par in range(len(X[bi])))
        lamD = min(_lam_min(Y[bi][par], dY[bi][par]) for bi in range(len(blocks)) for par in range(len(Y[bi])))
        alphaP = mp.mpf(1) if lamP > -gm else -gm / lamP
        alphaD = mp.mpf(1) if lamD > -gm else -gm / lamD
        if pd_feas:
            alphaP = alphaD = min(alphaP, alphaD)
        log.append(dict(iteration=it, mu=mu, pobj=pobj, dobj=dobj, gap=gap, primal_err_P=primal_err_P,
                        primal_err_p=primal_err_p, dual_err=dual_err, R_err=R_err, alphaP=alphaP, alphaD=alphaD,
                        beta=beta_c))
        for bi in range(len(blocks)):
            x[bi] = [a + alphaP * d for a, d in zip(x[bi], dx[bi])]
            X[bi] = [a + alphaP * d for a, d in zip(X[bi], dX[bi])]
            Y[bi] = [a + alphaD * d for a, d in zip(Y[bi], dY[bi])]
        y = [a + alphaD * d for a, d in zip(y, dy)]
