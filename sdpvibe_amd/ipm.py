"""SDPB's primal-dual interior-point method (Mehrotra predictor-corrector, SDPB 3.x) on the exact
multiprecision core: the counterpart of the reference's Solver (sdpvibe/solver.py:572-982).

The algorithm, its scalar decisions and its termination tests are SDPB's, as the reference states
them: residues (:721-764), termination (:798-816), Schur complement S and Q = sum P^T P with
P = L_S^-1 B (:818-869), predictor and corrector directions (:876-933), centering parameters and
step lengths (:935-951), the iterations.json record (:958-962). The organisation is this build's:

- per group of equally shaped blocks, one batched launch sequence; segment sums, the constraint
  matrix sum and the Schur complement's four-term products are single gathers over precomputed
  index maps instead of loops of sub-block copies;
- every reduction over all blocks is taken over all rows at once -- one exact GEMM for B^T x and
  P^T w, one `total` over the concatenated entries for c.x, Tr(XY) and <X + dX, Y + dY> -- and Q is
  accumulated exactly in residues with reference exponents equal to this iteration's column maxima
  (no flush, no chunk-order dependence). Hence the batched and the per-block ("plain") layouts of
  the same SDP produce bitwise identical iterates;
- the Cholesky factors of S and Q of the last iteration are kept (`final_factors`) for the adjoint.
"""
import os
import time

import mpmath as mp
import torch

from . import mpf as F
from .blockops import diagonal, keep, mirror_lower, parity_sum, repeat_batch, scale_columns, shift_diagonal, \
    fill_diagonal, stack_rows, twice
from .mpf import EZ, MPF, Chol, KSum, add, col_exps, gemm, max_abs, mul, symmetrize, total
from .steplen import min_eig, step_from_eig

DEFAULTS = dict(maxIterations=500, findPrimalFeasible=False, findDualFeasible=False,
                detectPrimalFeasibleJump=False, detectDualFeasibleJump=False,
                dualityGapThreshold="1e-30", primalErrorThreshold="1e-30", dualErrorThreshold="1e-30",
                initialMatrixScalePrimal="1e20", initialMatrixScaleDual="1e20",
                feasibleCenteringParameter="0.1", infeasibleCenteringParameter="0.3",
                stepLengthReduction="0.7", maxComplementarity="1e100", minPrimalStep="0", minDualStep="0")


def fmt_big(v, digits):
    """a number as SDPB prints a BigFloat: C++ defaultfloat with `digits` significant digits"""
    v = mp.mpf(v)
    if v == 0:
        return "0"
    s = mp.nstr(v, digits, min_fixed=-5, max_fixed=digits, strip_zeros=True)
    mant, _, ex = s.partition("e")
    if mant.endswith(".0"):
        mant = mant[:-2]
    if not ex:
        return mant
    sign = "-" if ex.startswith("-") else "+"
    ex = ex.lstrip("+-").lstrip("0") or "0"
    return f"{mant}e{sign}{ex.zfill(2)}"


def _diag_log2(ch):
    """log2 of the diagonal of the Cholesky factor L = D^-1 Lt (nb, n), from Lt's two top limbs"""
    d = torch.diagonal(ch.Lt.limbs, dim1=-2, dim2=-1)[:, :, :ch.n]          # (L, nb, n)
    top = d[-1].double() * 2.0 ** 32 + (d[-2].double() % 2.0 ** 32) if ch.L > 1 else d[-1].double()
    return torch.log2(torch.clamp(top, min=1e-300)) + (ch.Lt.exp + 32 * (ch.L - 2)) + ch.k.double()


def chol_log2_cond(ch, mask=None):
    """per block log2 of SDPB's Cholesky condition estimate (max / min diagonal of L)^2 over the rows
    selected by mask (nb, n) (all rows when None) (sdpvibe/solver.py:556-569)"""
    lg = _diag_log2(ch)
    if mask is None:
        mask = torch.ones_like(lg, dtype=torch.bool)
    hi = torch.where(mask, lg, torch.full_like(lg, -1e300)).amax(1)
    lo = torch.where(mask, lg, torch.full_like(lg, 1e300)).amin(1)
    return torch.where(mask.any(1), 2 * (hi - lo), torch.zeros_like(hi))


def chol_log2_cond_parts(ch):
    """(max, min) log2 of the diagonal of L for a single-matrix Chol"""
    lg = _diag_log2(ch)
    return float(lg.max()), float(lg.min())


def _flat(parts):
    """MPF list -> one flat MPF of all their entries (for an exact `total` over all of them)"""
    return MPF(torch.cat([p.m.reshape(p.L, -1) for p in parts], 1), torch.cat([p.e.reshape(-1) for p in parts]))


class Timers:
    """per-phase wall time (SDPB_GPU_TIMING=1; synchronises the device at every phase boundary)"""

    def __init__(self):
        self.on = os.environ.get("SDPB_GPU_TIMING") == "1"
        self.acc, self.t, self.cur = {}, None, None

    def __call__(self, name):
        if not self.on:
            return
        torch.cuda.synchronize()
        now = time.perf_counter()
        if self.cur is not None:
            self.acc[self.cur] = self.acc.get(self.cur, 0.0) + now - self.t
        self.t, self.cur = now, name

    def report(self):
        if not self.on:
            return {}
        self("end")
        return dict(self.acc)


class GraphedQ:
    """Chol(Q) and the one-vector solves with its factor as HIP graphs (graph.Captured), captured on
    first use and replayed in later iterations: Q and the right-hand side are copied into the
    captured buffers, results are cloned out of the graph's buffers. N is fixed for a solver, so the
    launch sequence is too; every replay is bitwise the eager computation. Attributes not defined
    here (Lt, k, n, not_pd, ...) are those of the captured factor."""

    def __init__(self, Q):
        from .graph import Captured
        self._fac = None
        self.Qs = MPF(Q.m.clone(), Q.e.clone())
        self._chol = Captured(lambda A: Chol(A, check=False), self.Qs)
        self._fac = self._chol.out
        self.rs = F.zeros((Q.shape[0], 1), Q.L)
        self._solves = {}

    def __getattr__(self, name):
        fac = self.__dict__.get("_fac")
        if fac is None:
            raise AttributeError(name)
        return getattr(fac, name)

    def factor(self, Q):
        self.Qs.m.copy_(Q.m)
        self.Qs.e.copy_(Q.e)
        self._chol.replay()
        if bool(self._fac.not_pd):
            raise ValueError("matrix is not positive definite (diagonal)")
        return self

    def _solve(self, which, v):
        if tuple(v.shape) != tuple(self.rs.shape):
            return getattr(self._fac, which)(v)
        if which not in self._solves:
            from .graph import Captured
            self._solves[which] = Captured(lambda r: getattr(self._fac, which)(r), self.rs)
        self.rs.m.copy_(v.m)
        self.rs.e.copy_(v.e)
        out = self._solves[which].replay()
        return MPF(out.m.clone(), out.e.clone())

    def solve_lower(self, v):
        return self._solve("solve_lower", v)

    def solve_upper(self, v):
        return self._solve("solve_upper", v)

    def solve(self, v):
        return self._solve("solve", v)


class _State:
    """per-group iterate: x (nb, P, 1), X and Y (2 nb, n, n)"""
    __slots__ = ("x", "X", "Y")

    def __init__(self, x, X, Y):
        self.x, self.X, self.Y = x, X, Y


class Solver:
    """Solver(sdp, L, **SDPB parameters).run(max_iter) -> result dict, with self.log the
    iterations.json records. `Solver.y`, `state` hold the iterate; `final_factors` the last
    iteration's Cholesky factors of X, Y, S and Q."""

    def __init__(self, sdp, L, precision=None, **params):
        self.sdp, self.L = sdp, L
        self.precision = precision or (32 * (L - 1) - 2)
        self.digits = int(mp.ceil(self.precision * mp.log10(2))) + 1
        p = dict(DEFAULTS)
        p.update(params)
        self.p = {k: (mp.mpf(v) if isinstance(v, str) else v) for k, v in p.items()}
        self.S = lambda v: F.scalar(v, L)
        self.state = []
        for g in sdp.groups:
            I = repeat_batch(F.eye(g.n, L, 1), g.nbp)
            X = fill_diagonal(I, self.S(self.p["initialMatrixScalePrimal"]), g.Xmask)
            Y = fill_diagonal(I, self.S(self.p["initialMatrixScaleDual"]), g.Xmask)
            self.state.append(_State(F.zeros((g.nb, g.P, 1), L), X, Y))
        self.y = F.zeros((sdp.N, 1), L)
        self._update = None                                   # (alphaP, alphaD, directions) pending
        self.log = []
        self.final_factors = None
        self.timers = Timers()
        self.iterations_done, self.alphaP, self.alphaD = 0, mp.mpf(0), mp.mpf(0)
        # chol(Q) and the solves with its factor replayed from HIP graphs after the first iteration
        self.graphs = os.environ.get("SDPB_GPU_GRAPHS", "1") != "0"
        self._qgraph = None
        self._b_exps = None                                  # column exponents of all of B (constant)

    # ---------------------------------------------------------------- streamed per-group arrays
    def _groups_prefetch(self, fetch):
        """(group, its index) in order, calling fetch(next group) before each group is used, so a
        streamed array of the next group is copied in while the current one is worked on"""
        groups = self.sdp.groups
        for i, g in enumerate(groups):
            if i + 1 < len(groups):
                fetch(groups[i + 1])
            yield i, g

    def _streamed(self, name, g):
        st = self.sdp.store
        return st is not None and st.is_host((name, g.gid))

    def _keep_P(self, g, w, P):
        if self._streamed("P", g):
            self.sdp.store.put(("P", g.gid), P)
        else:
            w["P"] = P

    def _P(self, g, w):
        """P = L_S^-1 B of the group on the device (resident, or copied in from host memory)"""
        return self.sdp.store.get(("P", g.gid)) if self._streamed("P", g) else w["P"]

    def _prefetch_P(self, g):
        if self._streamed("P", g):
            self.sdp.store.prefetch(("P", g.gid))

    def _B_exps(self):
        """row exponents of B^T over all blocks (B is constant: computed once)"""
        if self._b_exps is None:
            e = None
            for _, g in self._groups_prefetch(self.sdp.prefetch_B):
                t = col_exps(stack_rows(self.sdp.B_of(g)))
                e = t if e is None else torch.maximum(e, t)
            self._b_exps = e
        return self._b_exps

    # ---------------------------------------------------------------- block operators
    def constraint_sum(self, g, a):
        """sum_p A_p a_p per parity: (2 nb, d h, d h) from a (nb, P, 1) (sdpvibe/solver.py:610-621).
        All segments at once: V diag(a_s) V^T for every segment s as one batched product, off-diagonal
        segments halved, then scattered into the (r, c) and (c, r) blocks by one gather."""
        L, K, h, nbp = self.L, g.K, g.h, g.nbp
        nseg = g.P // K
        a2 = twice(a)                                                # (2nb, P, 1)
        aseg = MPF(a2.m.reshape(L, nbp * nseg, K, 1), a2.e.reshape(nbp * nseg, K, 1))
        Vr = MPF(g.V.m.unsqueeze(2).expand(-1, -1, nseg, -1, -1).reshape(L, nbp * nseg, h, K).contiguous(),
                 g.V.e.unsqueeze(1).expand(-1, nseg, -1, -1).reshape(nbp * nseg, h, K).contiguous())
        s = gemm(Vr, scale_columns(Vr, aseg), transB=True)         # (2nb nseg, h, h)
        if g.dim > 1:
            half = g.offdiag_seg.repeat(nbp)[:, None, None].expand(-1, h, h)
            s = s.shift(torch.where(half, -1, 0))
        sm = s.m.reshape(L, nbp, nseg * h * h)
        se = s.e.reshape(nbp, nseg * h * h)
        src = g.cm_seg * (h * h) + g.cm_ij
        n = g.n
        return MPF(sm[:, :, src].reshape(L, nbp, n, n).contiguous(), se[:, src].reshape(nbp, n, n).contiguous())

    def segment_trace(self, g, M):
        """sum over parities of the diagonal of the (r, c) K x K block of M (2 nb, d K, d K), as the
        vector over p = s K + k (nb, P, 1) (sdpvibe/solver.py:653-659)"""
        v = MPF(M.m[:, :, g.diag_rows, g.diag_cols].unsqueeze(-1).contiguous(),
                M.e[:, g.diag_rows, g.diag_cols].unsqueeze(-1).contiguous())
        return parity_sum(v, g.nb)

    def schur(self, g, AXI, AY):
        """S_pq = Tr(A_p X^-1 A_q Y) from AXI = Vx^T X^-1 Vx and AY = Vx^T Y Vx (sdpvibe/solver.py:623-651):
        dim 1: the parity sum of AXI o AY; dim > 1: four Hadamard products per parity gathered at once,
        summed in the reference's order, the parity sum quartered, the lower triangle mirrored"""
        L, nb, P = self.L, g.nb, g.P
        if g.dim == 1:
            S = parity_sum(mul(AXI, AY), nb)
        else:
            ax = AXI.m.reshape(L, g.nbp, -1)
            axe = AXI.e.reshape(g.nbp, -1)
            ay = AY.m.reshape(L, g.nbp, -1)
            aye = AY.e.reshape(g.nbp, -1)
            el = None
            for t in range(4):
                ia, iy = g.schur_axi[t], g.schur_ay[t]
                term = mul(MPF(ax[:, :, ia].contiguous(), axe[:, ia].contiguous()),
                           MPF(ay[:, :, iy].contiguous(), aye[:, iy].contiguous()))
                el = term if el is None else add(el, term)
            S = parity_sum(el, nb).shift(-2).reshape(nb, P, P)
        return mirror_lower(S)

    # ---------------------------------------------------------------- one iteration's phases
    def _apply_update(self):
        if self._update is None:
            return
        aP, aD, dirs = self._update
        for st, (dx, dX, dY) in zip(self.state, dirs):
            st.x = add(st.x, mul(dx, aP))
            st.X = add(st.X, mul(dX, aP))
            st.Y = add(st.Y, mul(dY, aD))
        self._update = None

    def _residues(self):
        """Cholesky factors of X and Y, pairings, residues and the iteration's scalars"""
        sdp, L = self.sdp, self.L
        out = dict(cond=(mp.mpf(0), ""), work=[])
        cx_terms, diag_terms = [], []
        dual_err, primal_err_P = mp.mpf(0), mp.mpf(0)
        best = -1.0
        xs = [stack_rows(st.x) for st in self.state]
        Bx = KSum(self._B_exps(), torch.stack([col_exps(x) for x in xs]).amax(0), L)   # B^T x over all blocks
        for gi, g in self._groups_prefetch(sdp.prefetch_B):
            st = self.state[gi]
            Bg = sdp.B_of(g)
            cX, cY = Chol(st.X), Chol(st.Y)
            for tag, ch in (("X", cX), ("Y", cY)):
                lg = chol_log2_cond(ch, g.Xmask).tolist()
                z = max(range(len(lg)), key=lambda i: lg[i])
                if lg[z] > best:
                    best = lg[z]
                    out["cond"] = (mp.mpf(2) ** mp.mpf(lg[z]), f"{tag}_cholesky.block_{g.idx[z % g.nb]}_{z // g.nb}")
            T = cX.solve_lower(g.Vx)
            AXI = mirror_lower(gemm(T, T, transA=True))
            AY = mirror_lower(gemm(g.Vx, gemm(st.Y, g.Vx), transA=True))
            By = gemm(Bg, repeat_batch(self.y, g.nb))
            Bx.add(stack_rows(Bg), xs[gi])
            del Bg
            dres = add(add(-self.segment_trace(g, AY), By, -1), g.c)
            Pres = keep(add(self.constraint_sum(g, st.x), st.X, -1), g.mask2)
            mXY = keep(-gemm(st.X, st.Y), g.mask2)
            dual_err = max(dual_err, max_abs(dres))
            primal_err_P = max(primal_err_P, max_abs(Pres))
            cx_terms.append(mul(g.c, st.x))
            diag_terms.append(diagonal(mXY))
            out["work"].append(dict(cX=cX, cY=cY, AXI=AXI, AY=AY, dres=dres, Pres=Pres, mXY=mXY))
        out["pobj"] = sdp.const + total(_flat(cx_terms))
        out["dobj"] = sdp.const + total(mul(sdp.b, self.y))
        out["pvec"] = add(sdp.b, Bx.result(), -1)
        out["mu"] = -total(_flat(diag_terms)) / sdp.total_psd_rows
        out["dual_err"], out["primal_err_P"] = dual_err, primal_err_P
        out["primal_err_p"] = max_abs(out["pvec"])
        return out

    def _schur_and_q(self, work):
        """Chol(S) per group, P = L_S^-1 B, Q = sum P^T P (exact, residues), Chol(Q)"""
        sdp, L = self.sdp, self.L
        TM = self.timers
        best, cond = -1.0, (mp.mpf(0), "")
        tops = None
        for gi, g in self._groups_prefetch(sdp.prefetch_B):
            w = work[gi]
            TM("schur: assemble S")
            S = self.schur(g, w.pop("AXI"), w.pop("AY"))
            TM("schur: chol(S)")
            cS = Chol(S)
            del S
            lg = chol_log2_cond(cS).tolist()
            q = max(range(len(lg)), key=lambda i: lg[i])
            if lg[q] > best:
                best = lg[q]
                cond = (mp.mpf(2) ** mp.mpf(lg[q]), f"schur_complement_cholesky.block_{g.idx[q]}")
            w["cS"] = cS
            TM("schur: P = L^-1 B")
            P = cS.solve_lower(sdp.B_of(g))
            Ps = stack_rows(P)
            t = F._top(Ps).amax(0)
            tops = t if tops is None else torch.maximum(tops, t)
            w["Pexp"] = col_exps(Ps)
            self._keep_P(g, w, P)
            del P, Ps
        TM("Q: accumulate")
        acc = F.GramAcc(sdp.N, L, tops)
        for gi, g in self._groups_prefetch(self._prefetch_P):
            acc.add(stack_rows(self._P(g, work[gi])))
        Q = mirror_lower(acc.result())
        TM("chol(Q)")
        if self.graphs and self.iterations_done >= 1:    # the first iteration runs eagerly (autotuner)
            if self._qgraph is None:
                self._qgraph = GraphedQ(Q)
            cQ = self._qgraph.factor(Q)
        else:
            cQ = Chol(Q)
        qc = mp.mpf(2) ** mp.mpf(float(chol_log2_cond(cQ)[0]))
        return cQ, qc, cond

    def _direction(self, work, pvec, cQ, mu, beta, corr, scal):
        """predictor (corr=False) or corrector direction (sdpvibe/solver.py:871-933): dy, and per group
        (dx, dX, dY); predictor: R-err and <X + dX, Y + dY>; corrector: the step-length eigenvalues"""
        sdp = self.sdp
        bm = self.S(beta * mu)
        for g, st, w in zip(sdp.groups, self.state, work):
            R = w["mXY"] if not corr else add(w["mXY"], gemm(w["dXp"], w["dYp"]), -1)
            R = shift_diagonal(R, bm, g.Xmask)
            Z = symmetrize(w["cX"].solve(add(gemm(w["Pres"], st.Y), R, -1)))
            VZV = gemm(g.Vx, gemm(Z, g.Vx), transA=True)
            rhs = add(-w["dres"], self.segment_trace(g, VZV), -1)
            w["R"], w["w"] = R, w["cS"].solve_lower(rhs)
        ws = [stack_rows(w["w"]) for w in work]
        Pw = KSum(torch.stack([w["Pexp"] for w in work]).amax(0), torch.stack([col_exps(v) for v in ws]).amax(0),
                  self.L)                                                  # P^T w over all blocks
        for gi, g in self._groups_prefetch(self._prefetch_P):
            Pw.add(stack_rows(self._P(g, work[gi])), ws[gi])
        if sdp.groups:
            self._prefetch_P(sdp.groups[0])
        dy = cQ.solve(add(pvec, Pw.result(), -1))
        del ws, Pw
        fr_terms, dirs = [], []
        lamP = lamD = None
        for gi, g in self._groups_prefetch(self._prefetch_P):
            st, w = self.state[gi], work[gi]
            Pdy = gemm(self._P(g, w), repeat_batch(dy, g.nb))
            dx = w["cS"].solve_upper(add(w["w"], Pdy))
            dX = add(self.constraint_sum(g, dx), w["Pres"])
            dY = keep(-symmetrize(w["cX"].solve(add(gemm(dX, st.Y), w["R"], -1))), g.mask2)
            if not corr:
                fr_terms.append(mul(keep(add(st.X, dX), g.mask2), keep(add(st.Y, dY), g.mask2)))
                w["dXp"], w["dYp"] = dX, dY
            else:
                gm = self.p["stepLengthReduction"]
                lamP = min_eig([w["cX"]], [dX], lam=lamP, gamma=gm, stats=scal)
                lamD = min_eig([w["cY"]], [dY], lam=lamD, gamma=gm, stats=scal)
                dirs.append((dx, dX, dY))
        if not corr:
            return dy, total(_flat(fr_terms)), None
        return dy, (lamP, lamD), dirs

    # ---------------------------------------------------------------- the iteration
    def run(self, max_iter=None, log=None, verbose=False, on_iteration=None, max_runtime=None,
            checkpoint=None, checkpoint_interval=None, stop=None):
        """iterate until one of SDPB's termination tests holds; returns the result record
        (terminateReason, objectives, gap, errors, iterations, time). on_iteration(row) receives each
        iterations.json record; checkpoint(solver) is called at the start of an iteration once
        checkpoint_interval seconds have passed; stop() -> True ends the run as SIGTERM does."""
        t0 = time.time()
        for row in self.iterate(max_iter, max_runtime, checkpoint, checkpoint_interval, stop):
            if log is not None:
                log.append(row)
            if on_iteration is not None:
                on_iteration(row)
            if verbose:
                last = self.last
                print(f"{row['iteration']:4d} {time.time() - t0:7.1f} mu={mp.nstr(last['mu'], 3)} "
                      f"P-obj={mp.nstr(last['pobj'], 8)} D-obj={mp.nstr(last['dobj'], 8)} gap={mp.nstr(last['gap'], 3)} "
                      f"steps={mp.nstr(last['alphaP'], 3)},{mp.nstr(last['alphaD'], 3)} beta={mp.nstr(last['beta'], 3)}",
                      flush=True)
        return self.result

    def iterate(self, max_iter=None, max_runtime=None, checkpoint=None, checkpoint_interval=None, stop=None):
        """generator of the iterations.json records, one per completed iteration; when it stops,
        self.result holds the verdict record"""
        p, sdp, L = self.p, self.sdp, self.L
        mp.mp.prec = F.NBITS(L) + 64
        fmt = lambda v: fmt_big(v, self.digits)
        alphaP, alphaD = self.alphaP, self.alphaD              # the last steps (0 before the first)
        t0 = time.time()
        last_ck = t0
        it = self.iterations_done                                # > 0 when resumed from a checkpoint
        pobj = dobj = gap = primal_err = dual_err = mp.mpf(0)
        self.timers = TM = Timers()

        def finish(reason, iterations):
            self.result = dict(terminateReason=reason, primalObjective=pobj, dualObjective=dobj, dualityGap=gap,
                               primalError=primal_err, dualError=dual_err, iterations=iterations,
                               time=time.time() - t0)

        while True:
            it += 1
            ti = time.time()
            if stop is not None and stop():
                finish("SIGTERM signal received", it - 1)
                return
            if checkpoint is not None and checkpoint_interval is not None and time.time() - last_ck >= checkpoint_interval:
                checkpoint(self)
                last_ck = time.time()
            TM("residues")
            self._apply_update()
            r = self._residues()
            pobj, dobj, mu = r["pobj"], r["dobj"], r["mu"]
            dual_err, primal_err_P, primal_err_p = r["dual_err"], r["primal_err_P"], r["primal_err_p"]
            primal_err = max(primal_err_P, primal_err_p)
            gap = abs(pobj - dobj) / max(abs(pobj) + abs(dobj), mp.mpf(1))
            dual_feas = dual_err < p["dualErrorThreshold"]
            primal_feas = primal_err < p["primalErrorThreshold"]
            pd_feas = dual_feas and primal_feas
            reason = None
            if pd_feas and gap < p["dualityGapThreshold"]:
                reason = "found primal-dual optimal solution"
            elif dual_feas and p["findDualFeasible"]:
                reason = "found dual feasible solution"
            elif primal_feas and p["findPrimalFeasible"]:
                reason = "found primal feasible solution"
            elif alphaD == 1 and p["detectDualFeasibleJump"]:
                reason = "dual feasible jump detected"
            elif alphaP == 1 and p["detectPrimalFeasibleJump"]:
                reason = "primal feasible jump detected"
            elif it > (max_iter or p["maxIterations"]):
                reason = "maxIterations exceeded"
            elif max_runtime is not None and time.time() - t0 >= max_runtime:
                reason = "maxRuntime exceeded"
            elif it > 1 and alphaP < p["minPrimalStep"]:
                reason = "primal step too small"
            elif it > 1 and alphaD < p["minDualStep"]:
                reason = "dual step too small"
            if reason is None and mu > p["maxComplementarity"]:
                reason = "maxComplementarity exceeded"
            if reason:
                finish(reason, it - 1)
                return
            TM("schur + Q")
            work = r["work"]
            cQ, Q_cond, scond = self._schur_and_q(work)
            cond = max(r["cond"], scond, key=lambda c: c[0])
            TM("predictor")
            scal = {}
            beta_p = mp.mpf(0) if pd_feas else p["infeasibleCenteringParameter"]
            R_err = max(max_abs(shift_diagonal(w["mXY"], self.S(mu), g.Xmask)) for g, w in zip(sdp.groups, work))
            _, fr, _ = self._direction(work, r["pvec"], cQ, mu, beta_p, False, scal)
            rr = fr / (mu * sdp.total_psd_rows)
            beta = rr * rr if rr < 1 else rr
            beta_c = min(max(p["feasibleCenteringParameter"], beta), mp.mpf(1)) if pd_feas \
                else max(p["infeasibleCenteringParameter"], beta)
            TM("corrector")
            dy, (lamP, lamD), dirs = self._direction(work, r["pvec"], cQ, mu, beta_c, True, scal)
            gm = p["stepLengthReduction"]
            alphaP, alphaD = step_from_eig(lamP, gm), step_from_eig(lamD, gm)
            if pd_feas:
                alphaP = alphaD = min(alphaP, alphaD)
            TM("update")
            aP, aD = self.S(alphaP), self.S(alphaD)
            self._update = (aP, aD, dirs)
            self.iterations_done, self.alphaP, self.alphaD = it, alphaP, alphaD
            self.y = add(self.y, mul(dy, aD))
            self.final_factors = dict(X=[w["cX"] for w in work], Y=[w["cY"] for w in work],
                                      S=[w["cS"] for w in work], Q=cQ)
            row = dict(iteration=it, total_time=round(time.time() - t0, 3), iter_time=round(time.time() - ti, 3),
                       mu=fmt(mu), **{"P-obj": fmt(pobj), "D-obj": fmt(dobj), "gap": fmt(gap),
                                      "P-err": fmt(primal_err_P), "p-err": fmt(primal_err_p), "D-err": fmt(dual_err),
                                      "R-err": fmt(R_err), "P-step": fmt(alphaP), "D-step": fmt(alphaD),
                                      "beta": fmt(beta_c), "Q_cond_number": fmt(Q_cond),
                                      "max_block_cond_number": fmt(cond[0]), "block_name": cond[1]})
            self.last = dict(mu=mu, pobj=pobj, dobj=dobj, gap=gap, primal_err_P=primal_err_P,
                             primal_err_p=primal_err_p, dual_err=dual_err, R_err=R_err, alphaP=alphaP,
                             alphaD=alphaD, beta=beta_c, eig_refined=scal.get("eig_refined", 0))
            self.log.append(row)
            yield row
