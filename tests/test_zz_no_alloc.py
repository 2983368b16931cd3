"""The allocation-free eigensolver entry point sdpv_symeig_static (its work area in the library's
device array) against sdpv_symeig with a caller-allocated work area, called through the C ABI: two
calls in a row, the second smaller, must give bitwise identical eigenvalues, eigenvectors, norms and
off-diagonal residues (same kernel, same arithmetic)."""
import numpy as np
import pytest
import torch

from sdpvibe_amd._lib import HipLibraryError, call, ptr, stream

pytestmark = pytest.mark.gpu


def _outputs(nb, n):
    z = lambda *s: torch.full(s, float("nan"), dtype=torch.float64, device="cuda")
    return z(nb, n), z(nb, n, n), z(nb), z(nb)


def _run_both(A):
    nb, n, _ = A.shape
    m = n + (n & 1)
    work = torch.empty((2 * nb * m * m,), dtype=torch.float64, device="cuda")
    ref = _outputs(nb, n)
    call("sdpv_symeig", ptr(A), nb, n, ptr(work), *(ptr(t) for t in ref), stream())
    got = _outputs(nb, n)
    call("sdpv_symeig_static", ptr(A), nb, n, *(ptr(t) for t in got), stream())
    torch.cuda.synchronize()
    return ref, got


def test_static_matches_allocating_entry():
    rng = np.random.default_rng(7)
    for nb, n in ((6, 57), (3, 20)):                       # second call smaller: stale contents must not matter
        A = torch.from_numpy(rng.standard_normal((nb, n, n))).cuda()
        ref, got = _run_both(A)
        for r, g in zip(ref, got):
            assert torch.equal(r, g)
        w = got[0].cpu().numpy()
        As = A.cpu().numpy()
        As = (As + As.transpose(0, 2, 1)) / 2
        assert np.allclose(w, np.linalg.eigvalsh(As), rtol=0, atol=1e-11 * np.abs(As).max())


def test_static_rejects_oversize():
    A = torch.zeros((1, 1, 1), dtype=torch.float64, device="cuda")
    out = _outputs(1, 1)
    with pytest.raises(HipLibraryError, match="unsupported size"):
        call("sdpv_symeig_static", ptr(A), 1 << 20, 3000, *(ptr(t) for t in out), stream())
