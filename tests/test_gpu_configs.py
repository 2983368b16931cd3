"""The benchmark configurations' Cholesky, solve and Gram-accumulation halves on the GPU, bit for bit
against the oracle's restatement of the same recursions (pinned in test_chol_oracle.py and
test_gram_oracle.py), with the error of the result stated in bits:

- the plain-kernel configuration: A = B B^T + 512 I at 768 bits (L = 24), B Gaussian MPF, seed 0;
  Cholesky and solves with 512 and with 1 right-hand side (sdpvibe/mpf.py:461-532);
- uneven SPD blocks A = D (B B^T + n I) D, D = diag(10^u), u in [-20, 20], n in 8..96, through the
  bucketed batched path (a few blocks of every size class against the oracle, all of them by residual);
- the Schur-complement chain at 1024 bits (L = 32): S_b of dimension 20..80, P_b = L_b^-1 B_b,
  Q = sum P_b^T P_b in residues (GramAcc), Chol(Q) and one solve (sdpvibe/solver.py:822-868,
  sdpvibe/mpf.py:536-630), at a scaled size against the oracle and at full size by properties.
"""
import numpy as np
import pytest
import torch

from oracle import batched_ref as BR
from oracle import gram_ref as GR
from oracle import oracle as O
from sdpvibe_amd import mpf as F
from sdpvibe_amd import synthetic as SY
from tests.intutil import from_planes

pytestmark = pytest.mark.gpu
EZ = F.EZ


def host(X):
    return X.m.cpu().numpy(), X.e.cpu().numpy()


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def _top_max(m, e):
    """max over entries of e + bitlen(m) (log2 of the largest magnitude, within one bit)"""
    L = m.shape[0]
    return int(O.mpf_top(m.reshape(L, -1), e.reshape(-1)).max())


def residual_bits(Am, Ae, Xm, Xe, Rm, Re):
    """log2 of max|R - A X| / (max|A| max|X|) for batched (L, nb, n, n), (L, nb, n, r), computed with
    the oracle's product (each entry rounded once at NB bits, far below the residual)"""
    L = Am.shape[0]
    Pm, Pe = O.mpf_gemm(Am, Ae, Xm, Xe)
    Dm, De = O.mpf_add(Rm.reshape(L, -1), Re.reshape(-1), Pm.reshape(L, -1), Pe.reshape(-1), -1)
    d = _top_max(Dm, De)
    if d <= EZ // 2:
        return -10 ** 6
    return d - _top_max(Am, Ae) - _top_max(Xm, Xe)


def factor_backward_bits(C, Am, Ae):
    """log2 of max |Lt Lt^T 2^-2F - D A D| for one factor (exact integers; D A D has unit-order
    diagonal): the backward error of the equilibrated factorisation"""
    L, n = C.L, C.n
    Fb = 32 * L - 40
    Lt = C.Lt.limbs.cpu().numpy()[:, 0, :n, :n]
    W = O.gemm_exact(Lt, Lt, 2 * L + 1)                             # exact Lt Lt^T
    llt = np.array(from_planes(W.reshape(2 * L + 1, -1)), dtype=object).reshape(n, n)
    a = np.array(from_planes(Am.reshape(L, -1)), dtype=object).reshape(n, n)
    k = C.k.cpu().numpy().reshape(n)
    e = Ae.reshape(n, n)
    worst = 0
    for i in range(n):
        for j in range(i + 1):
            if a[i, j] == 0:
                diff = abs(llt[i, j])
            else:
                sh = int(e[i, j]) - int(k[i]) - int(k[j]) + 2 * Fb      # D A D in units of 2^-2F
                v = a[i, j] << sh if sh >= 0 else None
                if v is None:                                           # not an integer at 2^-2F
                    diff = abs((llt[i, j] << -sh) - a[i, j]) >> -sh
                else:
                    diff = abs(llt[i, j] - v)
            worst = max(worst, diff)
    return worst.bit_length() - 2 * Fb if worst else -10 ** 6


# ------------------------------------------------------------------ plain-kernel configuration
@pytest.mark.timeout(900)
def test_plain_config_cholesky_and_solves():
    n, L = 512, 24
    B, A = SY.plain_config(n, L, seed=0)
    Bm, Be = SY.gaussian_mpf((n, n), L, seed=0)
    Gm, Ge = O.mpf_gemm(Bm[:, None], Be[None], Bm[:, None], Be[None], transB=True)
    Im = np.zeros((L, n, n), np.int32)
    Ie = np.full((n, n), EZ, np.int64)
    one = F.from_ints([n], [0], L, (1,))
    d = np.arange(n)
    Im[:, d, d] = one.m.cpu().numpy()[:, :1]
    Ie[d, d] = int(one.e[0])
    Am, Ae = O.mpf_add(Gm.reshape(L, -1), Ge.reshape(-1), Im.reshape(L, -1), Ie.reshape(-1))
    Am, Ae = Am.reshape(L, 1, n, n), Ae.reshape(1, n, n)
    gm, ge = host(A)
    _eq(gm, Am[:, 0], "A = B B^T + n I limbs")
    _eq(ge, Ae[0], "A = B B^T + n I exponents")

    C = F.Chol(A)
    ref = BR.Chol(Am, Ae)
    _eq(C.k.cpu().numpy().reshape(1, n), ref.k, "k")
    _eq(C.Lt.limbs.cpu().numpy(), ref.Lt.limbs, "Lt")
    bwd = factor_backward_bits(C, Am[:, 0], Ae[0])
    Fb = 32 * L - 40
    print(f"\n512x512 L=24 Cholesky: backward error 2^{bwd} (fixed point 2^-{Fb})")
    assert bwd <= -Fb + 8 + 9                                          # n 2^(8-F), n = 2^9

    for r, seed in ((512, 1), (1, 2)):
        Rm, Re = SY.gaussian_mpf((n, r), L, seed=seed)
        R = F.MPF(torch.from_numpy(Rm).cuda(), torch.from_numpy(Re).cuda())
        Y = C.solve_lower(R)
        X = C.solve_upper(Y)
        ym, ye = ref.solve_lower(Rm[:, None], Re[None])
        xm, xe = ref.solve_upper(ym, ye)
        _eq(Y.m.cpu().numpy(), ym[:, 0], f"solve_lower m ({r} RHS)")
        _eq(Y.e.cpu().numpy(), ye[0], f"solve_lower e ({r} RHS)")
        _eq(X.m.cpu().numpy(), xm[:, 0], f"solve_upper m ({r} RHS)")
        _eq(X.e.cpu().numpy(), xe[0], f"solve_upper e ({r} RHS)")
        rb = residual_bits(Am, Ae, xm, xe, Rm[:, None], Re[None])
        print(f"512x512 L=24 solve with {r} RHS: max|b - A x| / (max|A| max|x|) = 2^{rb}")
        assert rb <= -Fb + 8 + 18


# ------------------------------------------------------------------ uneven blocks, D = diag(10^u)
@pytest.mark.timeout(900)
def test_uneven_blocks_bucketed_vs_oracle():
    L = 24
    dims, blocks, rhs = SY.uneven_spd_blocks(400, L, seed=0)
    BC = F.BucketedChol(blocks)
    X = BC.solve(rhs)
    # a few blocks of every size class against the oracle, bit for bit
    classes = {}
    for i, n in enumerate(dims):
        classes.setdefault(-(-n // 32), []).append(i)
    picks = [i for c in sorted(classes) for i in classes[c][:3]]
    for i in picks:
        am, ae = host(blocks[i])
        bm, be = host(rhs[i])
        ref = BR.Chol(am[:, None], ae[None])
        ym, ye = ref.solve_lower(bm[:, None], be[None])
        xm, xe = ref.solve_upper(ym, ye)
        _eq(X[i].m.cpu().numpy(), xm[:, 0], f"block {i} (n={dims[i]}) m")
        _eq(X[i].e.cpu().numpy(), xe[0], f"block {i} (n={dims[i]}) e")
    # all 400 through the batched path: backward error of every solve, from the GPU's exact GEMM
    worst = -10 ** 6
    for A, x, b in zip(blocks, X, rhs):
        Rr = F.add(b, F.gemm(A, x), -1)
        t = F._top(Rr).max().item()
        if t > EZ // 2:
            worst = max(worst, t - F._top(A).max().item() - F._top(x).max().item())
    print(f"\n400 uneven blocks (cond up to ~1e80): max relative residual 2^{worst}")
    assert worst <= -(32 * L - 40) + 8 + 14


# ------------------------------------------------------------------ Schur-complement chain at L = 32
def _schur_chain(dims, S, B, N, r_seed):
    """GPU: P = L_b^-1 B_b per block (bucketed), Q = GramAcc(P), Chol(Q), x = Q^-1 r"""
    L = B.L
    BC = F.BucketedChol(S)
    offs = np.cumsum([0] + dims)
    Bs = [F.MPF(B.m[:, o:o + d].contiguous(), B.e[o:o + d].contiguous()) for o, d in zip(offs[:-1], dims)]
    P = BC.solve_lower(Bs)
    acc = F.GramAcc(N, L)
    acc.add(F.MPF(torch.cat([p.m for p in P], 1), torch.cat([p.e for p in P], 0)))
    Q = acc.result()
    from sdpvibe_amd.blockops import mirror_lower
    Qs = mirror_lower(Q)
    CQ = F.Chol(Qs)
    rm, re_ = SY.gaussian_mpf((N, 1), L, seed=r_seed)
    r = F.MPF(torch.from_numpy(rm).cuda(), torch.from_numpy(re_).cuda())
    x = CQ.solve(r)
    return P, acc, Q, Qs, CQ, r, x


@pytest.mark.timeout(900)
def test_schur_chain_scaled_vs_oracle():
    L, N = 32, 256
    dims, S, B = SY.schur_blocks(60, N, L, seed=0)
    P, acc, Q, Qs, CQ, r, x = _schur_chain(dims, S, B, N, r_seed=5)
    offs = np.cumsum([0] + dims)
    bm_all, be_all = host(B)
    pms, pes = [], []
    for b, (o, d) in enumerate(zip(offs[:-1], dims)):
        sm, se = host(S[b])
        ref = BR.Chol(sm[:, None], se[None])
        pm, pe = ref.solve_lower(bm_all[:, None, o:o + d], be_all[None, o:o + d])
        _eq(P[b].m.cpu().numpy(), pm[:, 0], f"P_{b} m")
        _eq(P[b].e.cpu().numpy(), pe[0], f"P_{b} e")
        pms.append(pm[:, 0])
        pes.append(pe[0])
    # the GPU adds all rows as one chunk (the reference exponents, hence the rounding, depend on
    # the chunking): so does the oracle
    G = GR.GramAcc(N, L)
    G.add(np.concatenate(pms, 1), np.concatenate(pes, 0))
    qm, qe = G.result()
    low = np.tril(np.ones((N, N), bool))
    gm, ge = host(Q)
    assert not np.argwhere((ge != qe) & low).size, "Q exponents"
    assert not np.argwhere((gm != qm).any(0) & low).size, "Q mantissas"
    assert acc.flushes == G.flushes
    sm, se = host(Qs)
    refQ = BR.Chol(sm[:, None], se[None])
    _eq(CQ.Lt.limbs.cpu().numpy(), refQ.Lt.limbs, "Chol(Q) Lt")
    rm, re_ = host(r)
    ym, ye = refQ.solve_lower(rm[:, None], re_[None])
    xm, xe = refQ.solve_upper(ym, ye)
    _eq(x.m.cpu().numpy(), xm[:, 0], "Q^-1 r m")
    _eq(x.e.cpu().numpy(), xe[0], "Q^-1 r e")


@pytest.mark.timeout(900)
def test_schur_chain_full_size_properties():
    """1200 blocks of dimension 20..80, N = 2000, L = 32: Q from GramAcc agrees with one exact GEMM over
    all rows to the working precision, and the solve of Q x = r has a small relative residual"""
    L, N = 32, 2000
    dims, S, B = SY.schur_blocks(1200, N, L, seed=1)
    P, acc, Q, Qs, CQ, r, x = _schur_chain(dims, S, B, N, r_seed=6)
    assert acc.flushes == 0
    Pall = F.MPF(torch.cat([p.m for p in P], 1), torch.cat([p.e for p in P], 0))
    del P, B, S
    torch.cuda.empty_cache()
    Qg = F.gemm(Pall, Pall, transA=True)
    low = torch.tril(torch.ones(N, N, dtype=torch.bool, device="cuda"))
    Ql = F.MPF(torch.where(low, Q.m, Qg.m), torch.where(low, Q.e, Qg.e))
    d = F.add(Ql, Qg, -1)
    td = F._top(d).max().item()
    rel = td - F._top(Qg).max().item() if td > EZ // 2 else -10 ** 6
    print(f"\nGramAcc vs one GEMM over {sum(dims)} rows: max|dQ| / max|Q| = 2^{rel}")
    assert rel <= -(32 * L - 2) + 16
    res = F.add(r, F.gemm(Qs, x), -1)
    tr = F._top(res).max().item()
    rb = tr - F._top(Qs).max().item() - F._top(x).max().item() if tr > EZ // 2 else -10 ** 6
    print(f"Chol(Q) solve, N={N}: max|r - Q x| / (max|Q| max|x|) = 2^{rb}")
    assert rb <= -(32 * L - 40) + 8 + 22
