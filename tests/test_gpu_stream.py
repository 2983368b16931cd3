"""GPU tests of the pieces that let an SDP stream through host memory (the reference's store.py
and Layout, sdpvibe/solver.py:227-398): the segment-wise exact sum KSum against one GEMM over the
stacked segments, the pinned host store, chunked layouts and streamed iterations against the
resident solver."""
import random

import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.synthetic import gaussian_mpf

pytestmark = pytest.mark.gpu


def ksum_case(put, L, M, N, sizes, seed, zero_col=True):
    """KSum over segments of the given row counts equals gemm(cat(A), cat(B), transA=True) bitwise"""
    rng = random.Random(seed)
    As, Bs = [], []
    for i, k in enumerate(sizes):
        am, ae = gaussian_mpf((k, M), L, seed=seed + 2 * i)
        bm, be = gaussian_mpf((k, N), L, seed=seed + 2 * i + 1)
        ae = ae + rng.randrange(-90, 90)                       # segments of different magnitudes
        if zero_col and M > 1:
            am[:, :, 1] = 0
            ae[:, 1] = F.EZ
        As.append(F.MPF(put(am), put(ae)))
        Bs.append(F.MPF(put(bm), put(be)))
    cat = lambda X: F.MPF(torch.cat([x.m for x in X], 1), torch.cat([x.e for x in X], 0))
    want = F.gemm(cat(As), cat(Bs), transA=True)
    EA = torch.stack([F.col_exps(a) for a in As]).amax(0)
    EB = torch.stack([F.col_exps(b) for b in Bs]).amax(0)
    acc = F.KSum(EA, EB, L)
    for a, b in zip(As, Bs):
        acc.add(a, b)
    got = acc.result()
    assert torch.equal(got.m, want.m) and torch.equal(got.e, want.e), (L, M, N, sizes)


@pytest.mark.parametrize("L,M,N,sizes", [(3, 40, 1, [7, 0, 30, 1]), (24, 300, 1, [270, 81, 1620]),
                                         (24, 64, 2, [5000, 100]), (9, 17, 5, [64, 65, 63])])
def test_ksum_equals_one_gemm_over_the_stack(L, M, N, sizes):
    ksum_case(lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda(), L, M, N, sizes, seed=L + M)


def test_store_round_trip_and_prefetch():
    """host keys come back bit for bit through page-locked memory, resident keys by reference; a
    second put of the same key reuses its buffer; a changed shape is refused"""
    from sdpvibe_amd.store import Store, nbytes
    am, ae = gaussian_mpf((3, 50, 7), 5, seed=1)
    A = F.MPF(torch.from_numpy(am).cuda(), torch.from_numpy(ae).cuda())
    tree = {"a": A, "l": [torch.arange(10, device="cuda"), (torch.ones(3, 3, device="cuda"), None)]}
    st = Store(host_keys={"h"})
    st.put("h", tree)
    st.put("d", A)
    assert st.get("d") is A
    want = [t.clone() for t in (A.m, A.e)]
    A.m.zero_()                                           # the stored copy must not alias the original
    got = st.get("h")
    assert torch.equal(got["a"].m, want[0]) and torch.equal(got["a"].e, want[1])
    assert torch.equal(got["l"][0], torch.arange(10, device="cuda")) and got["l"][1][1] is None
    n0 = st.host_bytes
    assert n0 >= nbytes(tree)
    st.put("h", got)                                      # same shapes: same buffer
    st.prefetch("h")
    again = st.get("h")
    assert st.host_bytes == n0 and torch.equal(again["a"].m, want[0])
    with pytest.raises(ValueError):
        st.put("h", {"a": F.MPF(A.m[:, :2], A.e[:2]), "l": got["l"]})
    assert st.bytes_h2d >= 2 * nbytes(tree) and st.bytes_d2h >= 2 * nbytes(tree)
    st.close()


@pytest.mark.parametrize("stream,chunk_gb", [("all", None), ("all", 1e-4), ("none", 1e-4)])
def test_placed_iterations_equal_resident(stream, chunk_gb, monkeypatch):
    """a memory plan that streams every B and P through host memory and/or splits the groups into
    small chunks gives the iterates of the all-resident batched layout bit for bit (five
    iterations: records, y, x, X)"""
    from sdpvibe_amd.ipm import Solver
    from sdpvibe_amd.layout import Layout
    from sdpvibe_amd.sdp import SDP
    sd = SDP.polynomial(n1=4, n3=2, degree=8, N=40, L=24, seed=5)
    monkeypatch.setenv("SDPB_GPU_STREAM", stream)
    lay = Layout(sd.shapes, sd.N, sd.L, chunk=None if chunk_gb is None else chunk_gb * 2 ** 30)
    if chunk_gb is not None:
        assert len(lay.chunks) > len(sd.groups)
    if stream == "all":
        assert len(lay.host_keys) == 2 * len(lay.chunks)
    sp = sd.placed(lay)
    A, B = Solver(sd, 24), Solver(sp, 24)
    A.run(max_iter=5)
    B.run(max_iter=5)
    assert len(A.log) == len(B.log) == 5
    for ra, rb in zip(A.log, B.log):
        for k in ("mu", "P-obj", "D-obj", "P-err", "p-err", "D-err", "R-err", "P-step", "D-step", "beta"):
            assert ra[k] == rb[k], (ra["iteration"], k)
    assert torch.equal(A.y.m, B.y.m) and torch.equal(A.y.e, B.y.e)
    def per_block(solver, name):
        """(even, odd) limbs of X or Y, or the limbs of x, per block in block order"""
        out = []
        for g, st in zip(solver.sdp.groups, solver.state):
            v = getattr(st, name).m
            for q in range(g.nb):
                out.append((v[:, q],) if name == "x" else (v[:, q], v[:, g.nb + q]))
        return out
    for name in ("x", "X", "Y"):
        for pa, pb in zip(per_block(A, name), per_block(B, name)):
            assert all(torch.equal(u, v) for u, v in zip(pa, pb)), name
    if stream == "all":
        assert sp.store.bytes_h2d > 0 and sp.store.bytes_d2h > 0
