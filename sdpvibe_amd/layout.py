"""Memory plan of an SDP on one GPU (the counterpart of Layout, sdpvibe/solver.py:227-398).

Blocks are grouped into chunks: runs of consecutive blocks of one shape (d, K, h), split so that a
chunk's working set fits the budget. Each chunk owns two large arrays of nb P N entries, its part
B of the constraint matrix and, during an iteration, P = L_S^-1 B; each either stays in HBM or
lives in page-locked host memory and is copied in when a phase needs it (store.Store), the next
chunk's copy overlapping the current chunk's work.

The plan never changes results: every reduction over blocks is exact over all rows at once (KSum,
GramAcc), so any chunking and any placement give the iterates of the all-resident batched layout
bit for bit. The budget may therefore follow the device's memory, unlike the reference, whose
chunks change rounding.

Controls (the reference's): SDPB_GPU_MEM_GB caps the budget (and the process's share of the
device memory), SDPB_GPU_CHUNK_GB caps a chunk's working set, SDPB_GPU_STREAM = auto (stream only
what does not fit), all (every B and P in host memory) or none (everything resident).
"""
import os
import sys

import torch

GB = 2 ** 30


class Layout:
    # working-set coefficients in MPF entries: Schur complement S and its factor (x nb P^2), the
    # triangular solve for P (x nb P N, besides P itself), pairings and their products
    # (x 2nb (dK)^2), X, Y, their factors and the direction pieces (x 2nb (dh)^2); Q, its factor
    # and the Gram accumulation (x N^2, plus the residue accumulator)
    A_S, A_B, A_A, A_X, A_Q = 9.0, 3.0, 5.0, 12.0, 7.0
    RESERVE = 1.5 * GB          # runtime context, library workspaces, fragmentation

    def __init__(self, shapes, N, L, budget=None, chunk=None, stream=None, n_primes=None):
        """shapes: (dim, K, h) per block, in block order"""
        self.shapes, self.N, self.L = list(shapes), N, L
        self.E = E = 4 * L + 8                               # bytes per MPF entry
        env_cap = float(os.environ.get("SDPB_GPU_MEM_GB", 0) or 0) * GB
        if env_cap and torch.cuda.is_available():        # emulate a smaller device: enforce the cap
            torch.cuda.set_per_process_memory_fraction(min(1.0, env_cap / self.device_memory()))
        if budget is None:
            budget = env_cap or self.device_memory() - self.RESERVE
        self.budget = budget
        chunk = float(os.environ.get("SDPB_GPU_CHUNK_GB", 0) or 0) * GB or chunk
        mode = os.environ.get("SDPB_GPU_STREAM", stream or "auto")
        if mode not in ("auto", "all", "none"):
            raise ValueError(f"SDPB_GPU_STREAM={mode}: one of auto, all, none")
        self.mode = mode
        nq = n_primes if n_primes is not None else self._primes(L + 1)
        self.fixed = (self.A_Q * E + 2 * nq) * N * N + self.RESERVE
        runs = self._runs()
        resident = self.fixed + sum(self.ws(c) + self.items(c)["B"] + self.items(c)["P"] for c in runs)
        self.host_keys = set()
        if chunk is None and (mode == "none" or (mode == "auto" and resident <= budget)):
            self.chunks = runs
            return
        avail = budget - self.fixed
        if avail <= 0:
            print(f"sdpvibe_amd: warning: Q and its factor need about {self.fixed / GB:.1f} GB (N = {N}), more than "
                  f"the GPU budget of {budget / GB:.1f} GB", file=sys.stderr)
        W = chunk or avail
        while True:
            self.chunks = self._split(runs, W)
            slot = max(self.items(c)["B"] + self.items(c)["P"] for c in self.chunks)   # the prefetched chunk
            need = max(self.ws(c) for c in self.chunks) + 2 * slot
            if chunk or W < 2 ** 26 or need <= avail or all(len(c) == 1 for c in self.chunks):
                break
            W *= 0.85
        if mode == "none":
            return
        if mode == "all":
            self.host_keys = {(name, gi) for gi in range(len(self.chunks)) for name in ("B", "P")}
            return
        # resident while they fit, most often copied first: P (five reads an iteration), then B (two)
        left = avail - need
        for name in ("P", "B"):
            for gi, c in enumerate(self.chunks):
                nbt = self.items(c)[name]
                if nbt <= left:
                    left -= nbt
                else:
                    self.host_keys.add((name, gi))

    @staticmethod
    def device_memory():
        return torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory

    @staticmethod
    def _primes(L):
        from .rns import PMAX, PMIN, _primes_desc
        bound = 4 * (1 << 17) * (1 << (32 * L - 1)) ** 2
        M, n = 1, 0
        for p in _primes_desc(PMAX, PMIN):
            if M > bound:
                break
            M *= p
            n += 1
        return n

    def _runs(self):
        runs, cur = [], []
        for j, s in enumerate(self.shapes):
            if cur and self.shapes[cur[0]] != s:
                runs.append(cur)
                cur = []
            cur.append(j)
        return runs + ([cur] if cur else [])

    def _split(self, runs, W):
        out = []
        for run in runs:
            cur = []
            for j in run:
                if cur and self.ws(cur + [j]) + self.items(cur + [j])["B"] + self.items(cur + [j])["P"] > W:
                    out.append(cur)
                    cur = []
                cur.append(j)
            out.append(cur)
        return out

    def shape(self, c):
        d, K, h = self.shapes[c[0]]
        return len(c), d, K, h, K * d * (d + 1) // 2

    def items(self, c):
        """bytes of the chunk's streamable arrays"""
        nb, d, K, h, P = self.shape(c)
        return {"B": nb * P * self.N * self.E, "P": nb * P * self.N * self.E}

    def ws(self, c):
        """bytes of the chunk's resident working set (everything but B and P)"""
        nb, d, K, h, P = self.shape(c)
        E = self.E
        return E * nb * (self.A_S * P * P + self.A_B * P * self.N) \
            + E * 2 * nb * (self.A_A * (d * K) ** 2 + self.A_X * (d * h) ** 2)

    def streamed(self):
        return bool(self.host_keys)

    def describe(self):
        nitems = 2 * len(self.chunks)
        hb = sum(self.items(self.chunks[gi])[n] for n, gi in self.host_keys)
        return (f"{len(self.chunks)} chunks, {nitems - len(self.host_keys)}/{nitems} arrays GPU-resident, "
                f"{hb / GB:.1f} GB streamed from host (GPU budget {self.budget / GB:.1f} GB)")
