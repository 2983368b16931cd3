"""SDP data for the interior-point driver: blocks of a polynomial-matrix program in SDPB's form
(sdpvibe/solver.py:1-13, :417-458):

  primal: minimise c.x + const  s.t.  X = sum_p A_p x_p >= 0,  B^T x = b
  dual  : maximise b.y + const  s.t.  Tr(A_p Y) + (B y)_p = c_p,  Y >= 0

Block j has dimension d, K sample points and bilinear bases V_even (he, K), V_odd (ho, K). Its
primal variables are indexed p = s K + k over the segments s = (c, r), c = 0..d-1, r = 0..c, and
the points k; A_p has v_k v_k^T (v_k column k of V, per parity) at the (r, c) and (c, r) blocks of
a (d h) x (d h) matrix, halved when r != c. Rows of c and B follow the same index.

Layout on the GPU. Blocks are gathered into groups of consecutive blocks with identical shapes
(d, K, h): one batched set of kernels per group ("batched"), or one group per block ("plain", the
per-block form). The even and odd parities of a group are stacked in one batch of 2 nb matrices
(even first). Both layouts keep the blocks in their order, so reductions over all blocks see the
same rows in the same order and the two layouts give bitwise identical iterates.
"""
import math
import time

import mpmath as mp
import numpy as np
import torch

from . import mpf as F
from .blockops import paste
from .mpf import EZ, MPF, add, gemm, mul
from .rns import dev


def segments(d):
    """(c, r, offset index s) of the d(d+1)/2 segments of a dimension-d block, in row order"""
    out = []
    for c in range(d):
        for r in range(c + 1):
            out.append((c, r, c * (c + 1) // 2 + r))
    return out


class Group:
    """nb consecutive blocks of one shape: dim d, K points, h basis rows (max of the parities)"""

    def __init__(self, idx, dim, K, h, L, N):
        self.idx, self.nb, self.dim, self.K, self.h = list(idx), len(idx), dim, K, h
        self.nbp = 2 * self.nb
        self.P = K * dim * (dim + 1) // 2
        self.n = dim * h                                    # size of the X, Y blocks
        self.L, self.N = L, N
        self._build_maps()

    def _build_maps(self):
        """index maps (on the device) of the segment gathers used by the driver"""
        d, K, h = self.dim, self.K, self.h
        segs = segments(d)
        # seg_diag: entry p = s K + k reads M[r K + k, c K + k] of a (d K) x (d K) matrix
        rows, cols = [], []
        for c, r, s in segs:
            for k in range(K):
                rows.append(r * K + k)
                cols.append(c * K + k)
        self.diag_rows = torch.tensor(rows, device=dev)
        self.diag_cols = torch.tensor(cols, device=dev)
        # cmws scatter: entry (i, j) of the (d h) x (d h) result takes segment (max, min) of the two
        # block indices, from s (i, j) when the row block is the smaller, else from s^T
        seg_of = {(c, r): s for c, r, s in segs}
        n = d * h
        src_seg = np.zeros((n, n), np.int64)
        src_ij = np.zeros((n, n), np.int64)
        for i in range(n):
            for j in range(n):
                bi, bj = i // h, j // h
                c, r = max(bi, bj), min(bi, bj)
                a, b = (i % h, j % h) if bi <= bj else (j % h, i % h)
                src_seg[i, j] = seg_of[(c, r)]
                src_ij[i, j] = a * h + b
        self.cm_seg = torch.from_numpy(src_seg.reshape(-1)).to(dev)
        self.cm_ij = torch.from_numpy(src_ij.reshape(-1)).to(dev)
        self.offdiag_seg = torch.tensor([c != r for c, r, _ in segs], device=dev)
        if d > 1:
            # Schur complement, (p0, p1) with p0 = (c0, r0, k0), p1 = (c1, r1, k1): the four terms
            # AXI[a, b] * AY[e, f] with (a, b, e, f) per term (sdpvibe/solver.py:634-650)
            P = self.P
            idx = np.zeros((4, 4, P, P), np.int64)
            for c0, r0, s0 in segs:
                for c1, r1, s1 in segs:
                    k0 = np.arange(K)[:, None]
                    k1 = np.arange(K)[None, :]
                    terms = (((c0, r1), (r0, c1)), ((r0, r1), (c0, c1)), ((c0, c1), (r0, r1)), ((r0, c1), (c0, r1)))
                    for t, ((ax, bx), (ay, by)) in enumerate(terms):
                        sl = (t, slice(None), slice(s0 * K, (s0 + 1) * K), slice(s1 * K, (s1 + 1) * K))
                        idx[sl[0], 0, sl[2], sl[3]] = ax * K + k0
                        idx[sl[0], 1, sl[2], sl[3]] = bx * K + k1
                        idx[sl[0], 2, sl[2], sl[3]] = ay * K + k0
                        idx[sl[0], 3, sl[2], sl[3]] = by * K + k1
            M = d * K
            self.schur_axi = torch.from_numpy((idx[:, 0] * M + idx[:, 1]).reshape(4, -1)).to(dev)
            self.schur_ay = torch.from_numpy((idx[:, 2] * M + idx[:, 3]).reshape(4, -1)).to(dev)


class SDP:
    """An SDP on the GPU: groups, per-group data (V, Vx, c, B, masks) and the objective (b, const).
    Construct with SDP.from_blocks (host block data), SDP.synthetic (random data of given block
    shapes) or SDP.polynomial (the strictly feasible polynomial-matrix program of the benchmark)."""

    def __init__(self, path, L, workers=None, budget=None, chunk=None, stream=None, layout="batched"):
        """the SDP stored at `path` (pmp2sdp output: directory or archive, binary or JSON blocks),
        placed by the memory plan for budget / chunk / stream (bytes, bytes, auto|all|none, as
        layout.Layout and the SDPB_GPU_* settings); `workers` is accepted for the reference's
        signature (blocks are decoded by the C++ codec in this process)"""
        self.__dict__.update(SDP.from_path(path, L, layout, budget=budget, chunk=chunk, stream=stream).__dict__)

    @classmethod
    def from_path(cls, path, L, layout="batched", budget=None, chunk=None, stream=None):
        """Blocks are decoded into host memory, the memory plan groups them into chunks, and the
        chunks are assembled on the device one after the other; the B of a chunk the plan places in
        host memory moves there as soon as it is assembled, so the SDP may exceed HBM."""
        from .layout import Layout
        from .sdp_io import SdpSource, read_block, read_normalization, read_objectives
        from .store import Store
        t0 = time.time()
        src = SdpSource(path)
        nblocks = src.json("control.json")["num_blocks"]
        const, b_strs = read_objectives(src)
        N = len(b_strs)
        raw, shapes, hs = [], [], []
        for j in range(nblocks):
            r = read_block(src, j, L)
            he, ho = r["bilinear_bases_even"][0], r["bilinear_bases_odd"][0]
            raw.append(r)
            shapes.append((r["dim"], r["K"], max(he, ho)))
            hs.append((he, ho))
        mp.mp.prec = F.NBITS(L) + 64
        t1 = time.time()
        self = cls._empty(L, N, mp.mpf(const), layout)
        if layout == "plain":
            self._plan(shapes)
        else:
            plan = Layout(shapes, N, L, budget=budget, chunk=chunk, stream=stream)
            self._set_groups(plan.chunks, shapes)
            self.layout_plan = plan
            self.store = Store(plan.host_keys) if plan.host_keys else None

        def on_device(j):
            r = raw[j]
            d, K = r["dim"], r["K"]
            P = K * d * (d + 1) // 2
            V = [None if arr is None else F.host_to_mpf(*arr, L, (h, K))
                 for h, arr in (r["bilinear_bases_even"], r["bilinear_bases_odd"])]
            return V, F.host_to_mpf(*r["c"], L, (P, 1)), F.host_to_mpf(*r["B"], L, (P, N))
        for g in self.groups:
            parts = [on_device(j) for j in g.idx]
            self._assemble(g, [p[0] for p in parts], [p[1] for p in parts], [p[2] for p in parts],
                           [hs[j] for j in g.idx])
            del parts
            if self.store is not None and self.store.is_host(("B", g.gid)):
                self.store.put(("B", g.gid), g.B)
                g.B = None
            for j in g.idx:
                raw[j] = None
        self._finish(F.from_strings(b_strs, L, (N, 1)))
        self.normalization = read_normalization(src)
        plan = self.layout_plan.describe() if self.layout_plan is not None else f"{len(self.groups)} groups, resident"
        self.read_info = f"read {t1 - t0:.1f}s, assemble {time.time() - t1:.1f}s; {plan}"
        return self

    # ---------------------------------------------------------------- construction
    @classmethod
    def _empty(cls, L, N, const, layout):
        self = cls.__new__(cls)
        mp.mp.prec = F.NBITS(L) + 64                        # the solver's scalar precision
        self.L, self.N, self.const, self.layout = L, N, mp.mpf(const), layout
        self.normalization = None
        self.store = None
        self.layout_plan = None
        self.read_info = ""
        return self

    def _plan(self, shapes):
        """groups of consecutive blocks with equal (dim, K, h); one per block in the plain layout"""
        groups, cur = [], []
        for j, s in enumerate(shapes):
            if cur and (self.layout == "plain" or shapes[cur[0]] != s):
                groups.append(cur)
                cur = []
            cur.append(j)
        if cur:
            groups.append(cur)
        self._set_groups(groups, shapes)

    def _set_groups(self, groups, shapes):
        self.groups = [Group(c, *shapes[c[0]], self.L, self.N) for c in groups]
        for gi, g in enumerate(self.groups):
            g.gid = gi
        self.shapes = shapes
        self.nblocks = len(shapes)

    # ---------------------------------------------------------------- placement (streaming)
    def B_of(self, g):
        """the group's B (nb, P, N) on the device: resident, or copied in from host memory"""
        return g.B if g.B is not None else self.store.get(("B", g.gid))

    def prefetch_B(self, g):
        if g.B is None:
            self.store.prefetch(("B", g.gid))

    def placed(self, layout=None):
        """this SDP grouped into the chunks of a memory plan (layout.Layout; by default the plan for
        this device and the SDPB_GPU_* settings), with the B of host-placed chunks moved to a
        store.Store in page-locked host memory. The iterates do not depend on the plan."""
        from .layout import Layout
        from .store import Store
        if layout is None:
            layout = Layout(self.shapes, self.N, self.L)
        new = _regroup(self, layout.chunks, "batched")
        new.layout_plan = layout
        new.store = Store(layout.host_keys) if layout.host_keys else None
        for g in new.groups:
            if new.store is not None and new.store.is_host(("B", g.gid)):
                new.store.put(("B", g.gid), g.B)
                g.B = None
        if new.store is not None:
            torch.cuda.synchronize()
        return new

    def _assemble(self, g, Vlist, clist, Blist, hlist):
        """per-group device arrays from per-block MPF pieces (V_even, V_odd (h_par, K)), c (P, 1), B (P, N)"""
        L, N, d, K, h, nb = self.L, self.N, g.dim, g.K, g.h, g.nb
        V = F.zeros((g.nbp, h, K), L)
        Xmask = torch.zeros((g.nbp, d * h), dtype=torch.bool, device=dev)
        for q in range(nb):
            for par in range(2):
                v = Vlist[q][par]
                z = par * nb + q
                hp = hlist[q][par]
                if hp:
                    V.m[:, z, :hp] = v.m
                    V.e[z, :hp] = v.e
                for c in range(d):
                    Xmask[z, c * h:c * h + hp] = True
        Vx = F.zeros((g.nbp, d * h, d * K), L)
        for c in range(d):
            paste(Vx, c * h, c * K, V)
        g.V, g.Vx, g.Xmask = V, Vx, Xmask
        g.mask2 = Xmask[:, :, None] & Xmask[:, None, :]
        g.c = MPF(torch.stack([x.m for x in clist], 1), torch.stack([x.e for x in clist], 0))
        g.B = MPF(torch.stack([x.m for x in Blist], 1), torch.stack([x.e for x in Blist], 0))
        g.psd_rows = int(Xmask.sum())

    def _finish(self, b):
        self.b = b
        self.total_psd_rows = sum(g.psd_rows for g in self.groups)

    @classmethod
    def from_blocks(cls, blocks, b, L, const=0, layout="batched"):
        """blocks: list of dicts (dim, K, Ve, Vo, c, B) with MPF values on the GPU (Ve (he, K) or None,
        Vo (ho, K) or None, c (P, 1), B (P, N)); b: MPF (N, 1)"""
        N = b.shape[0]
        self = cls._empty(L, N, const, layout)
        hs = [(0 if bl["Ve"] is None else bl["Ve"].shape[0], 0 if bl["Vo"] is None else bl["Vo"].shape[0])
              for bl in blocks]
        self._plan([(bl["dim"], bl["K"], max(h)) for bl, h in zip(blocks, hs)])
        for g in self.groups:
            self._assemble(g, [(blocks[j]["Ve"], blocks[j]["Vo"]) for j in g.idx], [blocks[j]["c"] for j in g.idx],
                           [blocks[j]["B"] for j in g.idx], [hs[j] for j in g.idx])
        self._finish(b)
        return self

    @classmethod
    def synthetic(cls, blocks, N, L, seed=0, layout="batched"):
        """random SDP with the given block shapes [dict(dim, K, he, ho)] (sdpvibe/solver.py:461-479),
        from a portable host generator: Gaussian MPF entries (synthetic.gaussian_mpf)"""
        from .synthetic import gaussian_mpf
        rng = np.random.default_rng(seed)

        def rnd(shape, s):
            m, e = gaussian_mpf(shape, L, seed=s, espread=4)
            return MPF(torch.from_numpy(m).to(dev), torch.from_numpy(e).to(dev))
        data = []
        for j, bl in enumerate(blocks):
            K, npair = bl["K"], bl["dim"] * (bl["dim"] + 1) // 2
            s = int(rng.integers(1 << 62))
            data.append(dict(dim=bl["dim"], K=K, Ve=rnd((bl["he"], K), s) if bl["he"] else None,
                             Vo=rnd((bl["ho"], K), s + 1) if bl["ho"] else None,
                             c=rnd((K * npair, 1), s + 2), B=rnd((K * npair, N), s + 3)))
        return cls.from_blocks(data, rnd((N, 1), int(rng.integers(1 << 62))), L, 0, layout)

    @classmethod
    def polynomial(cls, n1=300, n3=60, degree=26, N=3000, L=24, seed=0, coef_pow=8, layout="batched"):
        """The benchmark's polynomial-matrix program: n1 positivity constraints of dimension 1 and n3
        of dimension 3 (in that order), polynomials of the given degree in x >= 0, N free variables.

        - K = degree + 1 rescaled Laguerre sample points x_k = pi^2 (4k - 1)^2 / (-64 K log(3 - 2 sqrt 2)),
          sample scalings s_k = exp(-x_k);
        - bilinear bases orthonormal for the sampled measure: V_even[m, k] = sqrt(s_k) q_m(x_k) (degree
          <= degree/2), V_odd[m, k] = sqrt(x_k s_k) q'_m(x_k) (degree <= (degree-1)/2), rows
          orthonormal (V V^T = I), computed in mpmath and rounded to NB bits;
        - B rows: s_k sum_t a_t Lag_t(x_k) with Laguerre polynomials Lag_t, t <= degree, and
          coefficients a_t = g 10^u (g standard normal, u uniform in [-coef_pow, coef_pow]);
        - strictly feasible by construction: x* (1 on the diagonal segments, 0 elsewhere) gives
          X* = I, b = B^T x*; Y* = I and y* = 0.1 g give c = Tr(A_p) + B y*. The optimum is finite.
        Random numbers come from host PCG64 streams (seed), so the SDP is the same on every machine."""
        self = cls._empty(L, N, 0, layout)
        K = degree + 1
        he, ho = degree // 2 + 1, (degree - 1) // 2 + 1
        shapes = [(1, K, max(he, ho))] * n1 + [(3, K, max(he, ho))] * n3
        self._plan(shapes)
        Ve, Vo, Phi = _pmp_bases(K, he, ho, L)
        rng = np.random.default_rng(seed)
        ystar = F.MPF(*_f64_to_mpf(0.1 * rng.standard_normal(N), L)).reshape(N, 1)
        vnorm = _column_norms2(Ve, Vo)
        from .synthetic import powers_of_ten
        p10 = powers_of_ten(-coef_pow, coef_pow, L)
        for g in self.groups:
            d, nb = g.dim, g.nb
            npair = d * (d + 1) // 2
            # one random stream per block (seed, block index): the data do not depend on the grouping
            g_parts, u_parts = [], []
            for j in g.idx:
                gr = np.random.default_rng([seed, j])
                g_parts.append(gr.standard_normal(K * npair * N))
                u_parts.append(gr.integers(0, 2 * coef_pow + 1, K * npair * N))
            coef = _f64_to_mpf(np.concatenate(g_parts), L)
            u = torch.from_numpy(np.concatenate(u_parts)).to(dev)
            del g_parts, u_parts
            A = mul(MPF(*coef), MPF(p10.m[:, u].contiguous(), p10.e[u].contiguous()))
            del coef, u
            A = A.reshape(nb, K, npair * N)
            Bg = gemm(_repeat(Phi, nb), A)                           # (nb, K, npair N): rows k
            del A
            # (nb, k, s, n) -> (nb, s, k, n): rows p = s K + k
            Bm = Bg.m.reshape(L, nb, K, npair, N).permute(0, 1, 3, 2, 4).reshape(L, nb, npair * K, N).contiguous()
            Be = Bg.e.reshape(nb, K, npair, N).permute(0, 2, 1, 3).reshape(nb, npair * K, N).contiguous()
            del Bg
            Bblk = MPF(Bm, Be)
            # c = Tr(A_p) + B y*: Tr(A_p) = |ve_k|^2 + |vo_k|^2 on the diagonal segments
            tr = F.zeros((nb, npair * K, 1), L)
            for c, r, s in segments(d):
                if c == r:
                    tr.m[:, :, s * K:(s + 1) * K] = vnorm.m[:, None]
                    tr.e[:, s * K:(s + 1) * K] = vnorm.e[None]
            cg = add(tr, gemm(Bblk, _repeat(ystar, nb)))
            Vl = [(Ve, Vo)] * nb
            self._assemble(g, Vl, [MPF(cg.m[:, q], cg.e[q]) for q in range(nb)],
                           [MPF(Bm[:, q], Be[q]) for q in range(nb)], [(he, ho)] * nb)
            del Bm, Be, Bblk
        # b = B^T x*: the sum of the diagonal-segment rows of every block (one exact product)
        Brows, xs = [], []
        for g in self.groups:
            xst = torch.zeros((g.nb, g.P), dtype=torch.bool, device=dev)
            for c, r, s in segments(g.dim):
                if c == r:
                    xst[:, s * g.K:(s + 1) * g.K] = True
            Brows.append(g.B.reshape(g.nb * g.P, N))
            xs.append(xst.reshape(-1))
        Ball = MPF(torch.cat([x.m for x in Brows], 1), torch.cat([x.e for x in Brows], 0))
        one = F.scalar(1, L)
        sel = torch.cat(xs)
        xstar = MPF(torch.where(sel[None], one.m[:, :1], 0).reshape(L, -1, 1),
                    torch.where(sel, one.e[0], EZ).reshape(-1, 1))
        self._finish(gemm(Ball, xstar, transA=True))
        self.ystar = ystar
        return self


def _pick(X, idx, dim):
    """MPF rows idx (a list) along batch dimension dim of the exponent (dim + 1 of the limbs)"""
    t = torch.tensor(idx, device=X.e.device)
    return MPF(X.m.index_select(dim + 1, t), X.e.index_select(dim, t))


def _relayout(self, layout):
    """the same SDP grouped another way ("batched" or "plain"); data are copied, blocks keep their order"""
    plan = SDP._empty(self.L, self.N, self.const, layout)
    plan._plan(self.shapes)
    return _regroup(self, [g.idx for g in plan.groups], layout)


def _regroup(self, chunks, layout):
    """the same SDP with groups of the given consecutive equal-shape block lists"""
    new = SDP._empty(self.L, self.N, self.const, layout)
    new._set_groups(chunks, self.shapes)
    new.normalization = self.normalization
    where = {}
    for g in self.groups:
        for q, j in enumerate(g.idx):
            where[j] = (g, q)
    for ng in new.groups:
        srcs = [where[j] for j in ng.idx]
        # parity-stacked arrays: the even halves of all blocks, then the odd halves
        ng.V = _cat_batch([_pick(g.V, [q + half * g.nb], 0) for half in (0, 1) for g, q in srcs], 0)
        ng.Vx = _cat_batch([_pick(g.Vx, [q + half * g.nb], 0) for half in (0, 1) for g, q in srcs], 0)
        ng.Xmask = torch.cat([g.Xmask[q + half * g.nb][None] for half in (0, 1) for g, q in srcs], 0)
        ng.mask2 = ng.Xmask[:, :, None] & ng.Xmask[:, None, :]
        ng.c = _cat_batch([_pick(g.c, [q], 0) for g, q in srcs], 0)
        ng.B = _cat_batch([_pick(self.B_of(g), [q], 0) for g, q in srcs], 0)
        ng.psd_rows = int(ng.Xmask.sum())
    new._finish(self.b)
    return new


def _cat_batch(parts, dim):
    return MPF(torch.cat([p.m for p in parts], dim + 1), torch.cat([p.e for p in parts], dim))


SDP.relayout = _relayout


def _repeat(A, nb):
    return MPF(A.m.unsqueeze(1).expand(-1, nb, -1, -1).contiguous(), A.e.unsqueeze(0).expand(nb, -1, -1).contiguous())


def _f64_to_mpf(v, L):
    from .steplen import _mpf_from_float
    X = _mpf_from_float(torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev), L)
    return X.m, X.e


def _column_norms2(Ve, Vo):
    """(K, 1) MPF: |ve_k|^2 + |vo_k|^2 per point k, each square rounded once, summed in row order"""
    L = Ve.L
    K = Ve.shape[1]
    tot = None
    for V in (Ve, Vo):
        sq = mul(V, V)
        for m in range(V.shape[0]):
            row = MPF(sq.m[:, m].contiguous(), sq.e[m].contiguous()).reshape(K, 1)
            tot = row if tot is None else add(tot, row)
    return tot


def _pmp_bases(K, he, ho, L):
    """sample points, orthonormal bilinear bases and the Laguerre evaluation matrix, in mpmath at
    twice the working precision, rounded to NB bits: (V_even (he, K), V_odd (ho, K), Phi (K, K))"""
    old = mp.mp.prec
    mp.mp.prec = 2 * F.NBITS(L) + 64
    rho = 3 - 2 * mp.sqrt(2)
    xs = [mp.pi ** 2 * (4 * k - 1) ** 2 / (-64 * K * mp.log(rho)) for k in range(K)]
    sk = [mp.exp(-x) for x in xs]

    def orthonormal_rows(h, weight):
        rows = [[mp.sqrt(weight(k)) * xs[k] ** m for k in range(K)] for m in range(h)]
        out = []
        for r in rows:
            v = list(r)
            for _ in range(2):                                       # Gram-Schmidt, twice
                for q in out:
                    dp = mp.fsum(a * b for a, b in zip(v, q))
                    v = [a - dp * b for a, b in zip(v, q)]
            nrm = mp.sqrt(mp.fsum(a * a for a in v))
            out.append([a / nrm for a in v])
        return out

    ve = orthonormal_rows(he, lambda k: sk[k])
    vo = orthonormal_rows(ho, lambda k: xs[k] * sk[k])
    phi = [[sk[k] * mp.laguerre(t, 0, xs[k]) for t in range(K)] for k in range(K)]
    out = (F.from_mp([v for r in ve for v in r], L, (he, K)), F.from_mp([v for r in vo for v in r], L, (ho, K)),
           F.from_mp([v for r in phi for v in r], L, (K, K)))
    mp.mp.prec = old
    return out
