"""TEST INFRASTRUCTURE ONLY: the tiled Cholesky (sdpvibe/blockchol.py:32-94) restated on host arrays over
the oracle's primitives (mpf.gemm, mpf.add and the equilibrated Cholesky composition). MPF matrices
are pairs (m (L, R, C) int32, e (R, C) int64). Imported by tests/ only."""
import numpy as np

from . import batched_ref as BR
from . import oracle as O


def _T(A):
    return np.ascontiguousarray(A[0].transpose(0, 2, 1)), np.ascontiguousarray(A[1].T)


def gemm(A, B, transA=False, transB=False):
    m, e = O.mpf_gemm(A[0][:, None], A[1][None], B[0][:, None], B[1][None], transA, transB)
    return m[:, 0], e[0]


def add(A, B, sgn=1):
    L = A[0].shape[0]
    m, e = O.mpf_add(A[0].reshape(L, -1), A[1].reshape(-1), B[0].reshape(L, -1), B[1].reshape(-1), sgn)
    return m.reshape(A[0].shape), e.reshape(A[1].shape)


class Chol2:
    """mpf.Chol of one matrix (the 2-d interface)"""

    def __init__(self, A):
        self.c = BR.Chol(A[0][:, None], A[1][None])

    def solve_lower(self, B):
        m, e = self.c.solve_lower(B[0][:, None], B[1][None])
        return m[:, 0], e[0]

    def solve_upper(self, B):
        m, e = self.c.solve_upper(B[0][:, None], B[1][None])
        return m[:, 0], e[0]


def _sym_lower(A):
    m, e = A[0].copy(), A[1].copy()
    n = e.shape[0]
    iu = np.triu_indices(n, 1)
    m[:, iu[0], iu[1]] = A[0][:, iu[1], iu[0]]
    e[iu[0], iu[1]] = A[1][iu[1], iu[0]]
    return m, e


class BlockChol:
    def __init__(self, Q, b):
        m, e = Q
        N = e.shape[0]
        self.cuts = cuts = list(range(0, N, b)) + [N]
        T = self.T = len(cuts) - 1
        A = {}
        for i in range(T):
            for j in range(i + 1):
                A[i, j] = (np.ascontiguousarray(m[:, cuts[i]:cuts[i + 1], cuts[j]:cuts[j + 1]]),
                           np.ascontiguousarray(e[cuts[i]:cuts[i + 1], cuts[j]:cuts[j + 1]]))
        self.D = []
        for k in range(T):
            ck = Chol2(_sym_lower(A.pop((k, k))))
            self.D.append(ck)
            for i in range(k + 1, T):
                A[i, k] = _T(ck.solve_lower(_T(A[i, k])))
            for i in range(k + 1, T):
                for j in range(k + 1, i + 1):
                    A[i, j] = add(A[i, j], gemm(A[i, k], A[j, k], transB=True), -1)
        self.Lo = A

    def _rows(self, B, k):
        c0, c1 = self.cuts[k], self.cuts[k + 1]
        return np.ascontiguousarray(B[0][:, c0:c1]), np.ascontiguousarray(B[1][c0:c1])

    @staticmethod
    def _cat(parts):
        return np.concatenate([p[0] for p in parts], 1), np.concatenate([p[1] for p in parts], 0)

    def solve_lower(self, B):
        y = []
        for k in range(self.T):
            r = self._rows(B, k)
            for j in range(k):
                r = add(r, gemm(self.Lo[k, j], y[j]), -1)
            y.append(self.D[k].solve_lower(r))
        return self._cat(y)

    def solve_upper(self, Y):
        x = [None] * self.T
        for k in reversed(range(self.T)):
            r = self._rows(Y, k)
            for i in range(k + 1, self.T):
                r = add(r, gemm(self.Lo[i, k], x[i], transA=True), -1)
            x[k] = self.D[k].solve_upper(r)
        return self._cat(x)
