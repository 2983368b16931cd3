"""HIP graph capture of mpf.gemm: several replays, each with eager work and host synchronisation in
between, every replay bit-identical to the eager result on the same inputs."""
import numpy as np
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.graph import Captured
from sdpvibe_amd.synthetic import gaussian_mpf

pytestmark = pytest.mark.gpu


def _mpf(shape, L, seed):
    m, e = gaussian_mpf((1,) + shape, L, seed=seed)
    return F.MPF(torch.from_numpy(m).cuda(), torch.from_numpy(e).cuda())


@pytest.mark.parametrize("L,n", [(24, 128), (8, 300)])
def test_graph_replay_matches_eager(L, n):
    A, B = _mpf((n, n), L, 1), _mpf((n, n), L, 2)
    cap = Captured(F.gemm, A, B)
    for rep in range(4):
        # new inputs, copied into the captured buffers
        A2, B2 = _mpf((n, n), L, 10 + rep), _mpf((n, n), L, 20 + rep)
        for dst, src in ((A, A2), (B, B2)):
            dst.m.copy_(src.m)
            dst.e.copy_(src.e)
        C = cap.replay()
        other = F.gemm(_mpf((n, n), L, 99), B2)          # eager work between replays
        torch.cuda.synchronize()
        ref = F.gemm(A2, B2)
        assert torch.equal(C.m, ref.m) and torch.equal(C.e, ref.e), f"replay {rep}"
        del other
