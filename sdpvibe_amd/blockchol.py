"""Tiled right-looking Cholesky of a large N x N matrix with the interface of mpf.Chol (the
reference's sdpvibe/blockchol.py interface: BlockChol(holder, b), solve_lower/solve_upper/solve,
log2_cond).

The monolithic factorisation (mpf.Chol of the whole matrix) holds about 9 copies of Q. Here Q is cut
into b x b tiles (lower triangle) and factored tile column by tile column:
    L_kk = Chol(A_kk);  L_ik = A_ik L_kk^-T (i > k);  A_ij -= L_ik L_jk^T (i >= j > k)
so the peak is about Q/2 plus a few tiles. The rounding sequence is the reference's (one Chol per
diagonal tile, exact products rounded once, one MPF subtraction per update), but each step is
issued as few launches: the panel solves of a tile column are one solve of the column-concatenated
tiles (every column is aligned and rounded on its own, so this is bitwise the per-tile solve), and
the trailing updates of one step are batched GEMMs over all tile pairs of equal shape (batched
products round each block on its own, so this is bitwise the per-pair product).
"""
import torch

from .blockops import mirror_lower
from .mpf import MPF, Chol, add, gemm


def _cut(A, r0, r1, c0, c1):
    return MPF(A.m[:, r0:r1, c0:c1].contiguous(), A.e[r0:r1, c0:c1].contiguous())


def _stack(parts):
    """2-d MPF list of one shape -> batched (n, r, c)"""
    return MPF(torch.stack([p.m for p in parts], 1), torch.stack([p.e for p in parts], 0))


def _hcat(parts):
    return MPF(torch.cat([p.m for p in parts], 2), torch.cat([p.e for p in parts], 1))


def _vcat(parts):
    return MPF(torch.cat([p.m for p in parts], 1), torch.cat([p.e for p in parts], 0))


class BlockChol:
    """L L^T = Q by b x b tiles; `D[k]` the Chol of diagonal tile k, `Lo[i, k]` (i > k) the factor's
    off-diagonal tiles"""

    def __init__(self, holder, b=2048):
        """holder: list [Q] (MPF (N, N), lower triangle valid); Q is released once cut into tiles"""
        Q = holder.pop()
        N = Q.shape[0]
        self.N, self.b, self.L = N, b, Q.L
        self.cuts = cuts = list(range(0, N, b)) + [N]
        T = self.T = len(cuts) - 1
        tiles = {(i, j): _cut(Q, cuts[i], cuts[i + 1], cuts[j], cuts[j + 1]) for i in range(T) for j in range(i + 1)}
        del Q
        self.D = []
        for k in range(T):
            ck = Chol(mirror_lower(tiles.pop((k, k))))
            self.D.append(ck)
            below = list(range(k + 1, T))
            if not below:
                continue
            # panel: L_ik^T = L_kk^-1 A_ik^T for every i > k, as one solve of the concatenated columns
            sol = ck.solve_lower(_hcat([tiles[i, k].T for i in below]))
            c0 = 0
            for i in below:
                w = cuts[i + 1] - cuts[i]
                tiles[i, k] = _cut(sol, 0, sol.shape[0], c0, c0 + w).T
                c0 += w
            # trailing update A_ij -= L_ik L_jk^T, batched over pairs of equal shape
            pairs = {}
            for i in below:
                for j in range(k + 1, i + 1):
                    pairs.setdefault((tiles[i, k].shape, tiles[j, k].shape), []).append((i, j))
            for plist in pairs.values():
                prod = gemm(_stack([tiles[i, k] for i, _ in plist]), _stack([tiles[j, k] for _, j in plist]),
                            transB=True)
                new = add(_stack([tiles[i, j] for i, j in plist]), prod, -1)
                for q, (i, j) in enumerate(plist):
                    tiles[i, j] = MPF(new.m[:, q], new.e[q])
        self.Lo = tiles

    def _rows(self, B, k):
        return MPF(B.m[:, self.cuts[k]:self.cuts[k + 1]].contiguous(), B.e[self.cuts[k]:self.cuts[k + 1]].contiguous())

    def solve_lower(self, B):
        """L^-1 B, B: MPF (N, m) (forward substitution over the tile rows)"""
        y = []
        for k in range(self.T):
            r = self._rows(B, k)
            for j in range(k):
                r = add(r, gemm(self.Lo[k, j], y[j]), -1)
            y.append(self.D[k].solve_lower(r))
        return _vcat(y)

    def solve_upper(self, Y):
        """L^-T Y (backward substitution over the tile rows)"""
        x = [None] * self.T
        for k in range(self.T - 1, -1, -1):
            r = self._rows(Y, k)
            for i in range(k + 1, self.T):
                r = add(r, gemm(self.Lo[i, k], x[i], transA=True), -1)
            x[k] = self.D[k].solve_upper(r)
        return _vcat(x)

    def solve(self, B):
        return self.solve_upper(self.solve_lower(B))

    def log2_cond(self):
        """log2 of the Cholesky condition estimate (max / min diagonal of L)^2 over all tiles"""
        from .ipm import chol_log2_cond_parts
        parts = [chol_log2_cond_parts(ck) for ck in self.D]
        return 2 * (max(p[0] for p in parts) - min(p[1] for p in parts))
