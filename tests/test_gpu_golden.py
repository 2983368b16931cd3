"""The reference CPU baseline's own printed results as golden values (tests/golden/
mpf_baseline_check.json, extracted from cpu_baseline/cpu_*.txt of the reference): its generator
(cpu_baseline/mpf_baseline.c:34-56: a 32-bit LCG, entries a/b with a full-precision quotient,
A = G G^T + n I) restated here, factored with the GPU Cholesky and solved for the first column of
B. L[n-1][n-1] and (L^-1 B)[n-1][0] must match the printed 7 significant digits."""
import json
import os

import mpmath as mp
import pytest
import torch

from sdpvibe_amd import mpf as F
from sdpvibe_amd.blockops import shift_diagonal
from tests.intutil import norm_closed_form

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "mpf_baseline_check.json")))


def _lcg_values(count, s=1):
    """the baseline's rnd(): a = (s >> 8) - 2^23, b = 2 (s' >> 9) + 1, value a / b"""
    out = []
    for _ in range(count):
        s = (s * 1103515245 + 12345) & 0xFFFFFFFF
        a = (s >> 8) - 8388608
        s = (s * 1103515245 + 12345) & 0xFFFFFFFF
        b = 2 * (s >> 9) + 1
        out.append((a, b))
    return out, s


def _to_mpf_ints(vals, L):
    """a / b -> (mantissa, exponent) rounded half up on the magnitude to NB bits"""
    NB = 32 * L - 2
    ms, es = [], []
    for a, b in vals:
        if a == 0:
            ms.append(0)
            es.append(F.EZ)
            continue
        sh = NB + 2 + b.bit_length() - abs(a).bit_length()
        q = (abs(a) << sh) // b                       # >= NB + 1 bits; truncation below the rounding bit
        r, shift = norm_closed_form(q if a > 0 else -q, NB)
        ms.append(r)
        es.append(shift - sh)
    return ms, es


@pytest.mark.parametrize("case", CASES, ids=[c["file"] for c in CASES])
def test_mpf_baseline_check_values(case):
    n, R, prec = case["n"], case["nrhs"], case["prec"]
    L = (prec + 31) // 32
    g, s = _lcg_values(n * n)
    b, _ = _lcg_values(n * R, s)
    bcol = [b[i * R] for i in range(n)]                # column 0 of B (row-major n x R)
    gm, ge = _to_mpf_ints(g, L)
    G = F.from_ints(gm, ge, L, (n, n))
    bm, be = _to_mpf_ints(bcol, L)
    B = F.from_ints(bm, be, L, (n, 1))
    A = shift_diagonal(F.gemm(G, G, transB=True), F.scalar(n, L))
    ch = F.Chol(A)
    mp.mp.prec = 64
    lt = F.MPF(ch.Lt.limbs[:, 0, n - 1, n - 1].reshape(L, 1).contiguous(),
               torch.tensor([ch.Lt.exp], dtype=torch.int64, device="cuda"))
    Lnn = lt.to_mp()[0] * mp.mpf(2) ** int(ch.k[0, n - 1])
    P = ch.solve_lower(B)
    pn0 = P[n - 1:n, 0:1].to_mp()[0, 0]
    for got, want in ((Lnn, case["L_last"]), (pn0, case["P_last_0"])):
        w = mp.mpf(want)
        assert abs(got - w) <= abs(w) * mp.mpf("6e-7"), (mp.nstr(got, 12), want)
