"""TEST INFRASTRUCTURE ONLY: mpf.GramAcc (sdpvibe/mpf.py:536-630) restated on host arrays over the
oracle's primitives. Chunks are rounded to the running reference exponents exactly as the
reference does; the accumulated sum is the exact integer Gram matrix of all rounded chunks since
the last flush (one exact product over the concatenated rows), normalised once per flush.
Imported by tests/ only."""
import numpy as np

from . import oracle as O

EZ = O.EZ


class GramAcc:
    HEAD = 28

    def __init__(self, N, L, E_ref=None):
        self.N, self.L = N, L
        self.E = np.full(N, EZ, np.int64) if E_ref is None else np.array(E_ref, np.int64)
        self.Tmax = np.full(N, EZ, np.int64)
        self.chunks = []                                   # rounded (L + 1, N, rows) limbs since the flush
        self.part = None                                   # (m (L, N, N), e (N, N))
        self.flushes = 0

    def add(self, m, e):
        """m (L, rows, N), e (rows, N)"""
        L, N = self.L, self.N
        NB = 32 * L - 2
        tp = O.mpf_top(m.reshape(L, -1), e.reshape(-1)).reshape(e.shape).max(0)
        live = tp > EZ // 2
        self.Tmax = np.maximum(self.Tmax, tp)
        known = self.E > EZ // 2
        if (live & known & (tp > self.E + self.HEAD)).any():
            self._flush()
            self.E = np.where(live & (tp > self.E), tp, self.E)
            known = self.E > EZ // 2
        self.E = np.where(live & ~known, tp, self.E)
        mt = np.ascontiguousarray(m.transpose(0, 2, 1))    # (L, N, rows)
        et = np.ascontiguousarray(e.T)
        s = ((self.E - NB)[:, None] - et).reshape(-1)
        s = np.where(et.reshape(-1) == EZ, 1 << 30, s)
        X, _ = O.shift_round(mt.reshape(L, -1), s, L + 1)
        self.chunks.append(X.reshape(L + 1, N, -1))

    def _crt(self):
        L, N = self.L, self.N
        NB = 32 * L - 2
        X = np.concatenate(self.chunks, 2)
        rows = X.shape[2]
        nw = 2 * (L + 1) + 2 + (max(rows, 1).bit_length() + 31) // 32
        Wd = O.gemm_exact(X, X, nw)                         # (nw, N, N)
        rE = np.where(self.E > EZ // 2, self.E - NB, 0)
        ein = (rE[:, None] + rE[None, :]).reshape(-1)
        om, oe = O.mpf_norm(Wd.reshape(nw, -1), ein, L)
        return om.reshape(L, N, N), oe.reshape(N, N)

    def _flush(self):
        if not self.chunks:
            return
        q = self._crt()
        self.part = q if self.part is None else self._add(self.part, q)
        self.chunks = []
        self.flushes += 1

    def _add(self, a, b):
        L = self.L
        om, oe = O.mpf_add(a[0].reshape(L, -1), a[1].reshape(-1), b[0].reshape(L, -1), b[1].reshape(-1))
        return om.reshape(a[0].shape), oe.reshape(a[1].shape)

    def next_reference(self, margin=4):
        return np.where(self.Tmax > EZ // 2, self.Tmax + margin, self.Tmax)

    def precision_loss(self):
        ok = (self.Tmax > EZ // 2) & (self.E > EZ // 2)
        return int((self.E - self.Tmax)[ok].max()) if ok.any() else 0

    def result(self):
        if not self.chunks:
            return self.part
        q = self._crt()
        self.chunks = []
        return q if self.part is None else self._add(self.part, q)
