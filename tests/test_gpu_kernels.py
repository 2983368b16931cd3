"""GPU parity: every HIP entry point against the C oracle on the same seeded inputs (bit-exact).
Runs only on a gfx950 GPU with libsdpvibe_hip.so built."""
import random

import numpy as np
import pytest
import torch

from oracle import oracle as O
from sdpvibe_amd import _lib, limbops
from sdpvibe_amd import mpf as F
from sdpvibe_amd import mpmat
from sdpvibe_amd.synthetic import gaussian_mpf
from tests.intutil import EZ, from_planes, rand_mpf_ints, to_planes

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _eq(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def test_mfma_fragment_layout():
    rng = np.random.default_rng(0)
    A = rng.integers(-128, 128, (32, 32)).astype(np.int8)
    B = rng.integers(-128, 128, (32, 32)).astype(np.int8)
    A[0, :] = -128
    B[:, 0] = -128
    dA, dB = cu(A), cu(B)
    D = torch.empty((32, 32), dtype=torch.int32, device="cuda")
    _lib.call("sdpv_mfma_selftest", _lib.ptr(dA), _lib.ptr(dB), _lib.ptr(D), _lib.stream())
    _eq(host(D), A.astype(np.int64) @ B.astype(np.int64), "mfma")


def _rand_words(rng, L, n):
    lim = 1 << (32 * L - 1)
    vals = [0, 1, -1, lim - 1, -lim] + [rng.randrange(-lim, lim) >> rng.randrange(0, 32 * L) for _ in range(n - 5)]
    return vals


@pytest.mark.parametrize("Lin,Lout", [(1, 1), (3, 2), (5, 7), (24, 24), (24, 25), (33, 34)])
def test_shift_round(Lin, Lout):
    rng = random.Random(Lin * 100 + Lout)
    n = 3000
    X = to_planes(_rand_words(rng, Lin, n), Lin)
    s = np.array([rng.choice([0, 1, -1, 31, 32, 33, 1 << 30, -(1 << 20)]) if i % 5 == 0 else
                  rng.randrange(-40 * Lout, 40 * Lin) for i in range(n)], np.int64)
    ref, rflag = O.shift_round(X, s, Lout)
    dX, ds = cu(X), cu(s)
    out = torch.empty((Lout, n), dtype=torch.int32, device="cuda")
    flag = torch.zeros(n, dtype=torch.int32, device="cuda")
    _lib.call("sdpv_shift_round", _lib.ptr(dX), Lin, n, _lib.ptr(ds), _lib.ptr(out), Lout, _lib.ptr(flag), 1,
              _lib.stream())
    _eq(host(out), ref, "shift_round")
    _eq(host(flag), rflag, "overflow flags")
    _eq(host(limbops.shift_round(dX, ds, Lout)), ref, "shift_round wrapper")


@pytest.mark.parametrize("L", [1, 4, 24, 64])
def test_add_same_bitlen(L):
    rng = random.Random(L)
    a, b = _rand_words(rng, L, 2000), _rand_words(rng, L, 2000)
    A, B = to_planes(a, L), to_planes(b, L)
    for sgn in (1, -1):
        _eq(host(limbops.add_same(cu(A), cu(B), sgn)), O.add_same(A, B, sgn), f"add_same {sgn}")
    _eq(host(limbops.bitlen_words(cu(A))), O.bitlen(A), "bitlen")


@pytest.mark.parametrize("L", [1, 4, 28, 34])
@pytest.mark.parametrize("uniform", [False, True])
def test_badd_align(L, uniform):
    rng = random.Random(19 + L)
    nb, rows, cols = 3, 5, 10
    lim = 1 << (32 * L - 1)
    a = [rng.randrange(-lim, lim) for _ in range(nb * rows * cols)]
    b = list(reversed(a))
    a[:5] = [0, 1, -1, lim - 1, -lim]
    A = to_planes(a, L).reshape(L, nb, rows, cols)
    B = to_planes(b, L).reshape(L, nb, rows, cols)
    ea = np.full((nb, cols), -100, np.int64)
    eb = ea.copy() if uniform else ea + np.array([0, 1, 31, 32, 33, 63, 64, 65, -10000, 10000])[None, :]
    for sgn in (1, -1):
        ref, rEO = O.badd_align(A, B, ea, eb, sgn)
        out = torch.empty((L, nb, rows, cols), dtype=torch.int32, device="cuda")
        EO = torch.empty((nb, cols), dtype=torch.int64, device="cuda")
        dA, dB, dea, deb = cu(A), cu(B), cu(ea), cu(eb)
        _lib.call("sdpv_badd_align", _lib.ptr(dA), _lib.ptr(dB), _lib.ptr(dea), _lib.ptr(deb), _lib.ptr(out),
                  _lib.ptr(EO), L, nb, rows, cols, sgn, _lib.stream())
        _eq(host(out), ref, "badd")
        _eq(host(EO), rEO, "badd exps")


@pytest.mark.parametrize("L,LI", [(1, 3), (2, 2), (10, 24), (24, 54), (26, 160)])
def test_mpf_norm(L, LI):
    rng = random.Random(L * 7 + LI)
    vals = _rand_words(rng, LI, 3000)
    X = to_planes(vals, LI)
    ein = np.array([rng.randrange(-5000, 5000) for _ in vals], np.int64)
    rm, re_ = O.mpf_norm(X, ein, L)
    r = F.normalize(cu(X), cu(ein), L)
    _eq(host(r.m), rm, "norm m")
    _eq(host(r.e), re_, "norm e")


@pytest.mark.parametrize("L", [1, 2, 10, 24, 26, 34, 64])
def test_mpf_mul_add_top(L):
    rng = random.Random(71 + L)
    n = 2000
    a, ea = rand_mpf_ints(rng, n, L)
    b, eb = rand_mpf_ints(rng, n, L)
    b[30], eb[30] = -a[30], ea[30]
    A, B = to_planes(a, L), to_planes(b, L)
    ea, eb = np.array(ea, np.int64), np.array(eb, np.int64)
    ea[7] = EZ
    dA, dB = F.MPF(cu(A), cu(ea)), F.MPF(cu(B), cu(eb))
    rm, re_ = O.mpf_mul(A, ea, B, eb)
    r = F.mul(dA, dB)
    _eq(host(r.m), rm, "mul m")
    _eq(host(r.e), re_, "mul e")
    rs = F.mul(dA, F.MPF(cu(B[:, 3:4]), cu(eb[3:4])))
    rm, re_ = O.mpf_mul(A, ea, B[:, 3:4], eb[3:4])
    _eq(host(rs.m), rm, "mul scalar")
    _eq(host(rs.e), re_, "mul scalar e")
    for sgn in (1, -1):
        rm, re_ = O.mpf_add(A, ea, B, eb, sgn)
        r = F.add(dA, dB, sgn)
        _eq(host(r.m), rm, f"add m {sgn}")
        _eq(host(r.e), re_, f"add e {sgn}")
    _eq(host(F._top(dA)), O.mpf_top(A, ea), "top")


@pytest.mark.parametrize("L", [1, 10, 28, 34])
@pytest.mark.parametrize("K", [17, 65])
def test_residues(L, K):
    # tests/test_rns_kernels.py:28-52 shapes, against the oracle
    rng = random.Random(123 + L + K)
    lim = 1 << (32 * L - 1)
    rows = 3
    values = [[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(rows)]
    values[0][3:8] = [0, 1, -1, -lim, lim - 1]
    X = np.stack([to_planes(r, L) for r in values], axis=1)
    R = mpmat.ctx(L)
    groups = {(0, R.n), (max(0, R.n - 16), min(16, R.n)), (min(5, R.n - 1), min(40, R.n - min(5, R.n - 1)))}
    for g0, G in sorted(groups):
        hi, lo = R.residues(cu(X), 3, K - 5, g0, G)
        got = 128 * host(hi).astype(np.int32) + host(lo)
        ref = O.residues(X, 3, K - 5, np.array(R.mods[g0:g0 + G]))
        _eq(got[:, :, :K - 5], ref, f"residues g0={g0}")
        assert (got[:, :, K - 5:] == 0).all()


@pytest.mark.parametrize("L", [1, 10, 24, 34, 130])
@pytest.mark.parametrize("width", [1, 5, 11])
def test_crt_block(L, width):
    # tests/test_rns_kernels.py:55-80 (L=130: n > 512 primes, the wide digit path)
    R = mpmat.ctx(L)
    rng = random.Random(41)
    rows, N, start = 5, 13, 2
    limit = R.M // 4
    vals = [[rng.randrange(-limit + 1, limit) for _ in range(N)] for _ in range(rows)]
    vals[0][2:7] = [0, 1, -1, limit - 1, -limit + 1]
    vals[1][2:7] = [1 << b for b in (0, 31, 32, 63, 64)]
    vals[2][2:7] = [-(1 << b) for b in (0, 31, 32, 63, 64)]
    acc = np.array([[[v % p for v in row] for row in vals] for p in R.mods], np.int16)
    W, bl, Bq = R.crt_block(cu(acc), N, start, width, rows)
    assert Bq == rows * width
    exp = [v for row in vals for v in row[start:start + width]]
    sel = acc[:, :, start:start + width].reshape(R.n, -1)
    rW, rbl = O.crt(sel, np.array(R.mods), R.E // 4)
    _eq(host(W), rW, "crt words")
    _eq(host(bl), rbl, "crt bitlen")
    assert from_planes(host(W)) == exp
    # the same integers from the product's CRT-ready form y = centred(residue * (M/m)^-1 mod m)
    W2, bl2, _ = R.crt_block(cu(scaled_residues(R, acc)), N, start, width, rows, scaled=True)
    _eq(host(W2), rW, "crt words (scaled input)")
    _eq(host(bl2), rbl, "crt bitlen (scaled input)")


def scaled_residues(R, acc):
    m = np.array(R.mods, np.int64)[:, None, None]
    y = acc.astype(np.int64) * np.array([int(v) for v in R.inv32.cpu()], np.int64)[:, None, None] % m
    return np.where(y > m // 2, y - m, y).astype(np.int16)


def test_product_accumulate():
    """RNS.accumulate (three INT8 products + modular combine) against numpy int64"""
    L = 24
    R = mpmat.ctx(L)
    rng = np.random.default_rng(5)
    for (M, N, K) in ((7, 5, 30), (64, 64, 64), (130, 70, 200), (300, 257, 1000)):
        G = R.n
        m = np.array(R.mods, np.int64)[:, None, None]
        ra = rng.integers(-8128, 8129, (G, M, K))
        rb = rng.integers(-8128, 8129, (G, N, K))
        ra = np.where(np.abs(ra) <= (m - 1) // 2, ra, 0)
        rb = np.where(np.abs(rb) <= (m - 1) // 2, rb, 0)
        kp = R.kp(K)
        pad = lambda r: np.pad(r, ((0, 0), (0, 0), (0, kp - K)))
        sp = lambda r: (((r + 64) & 127) - 64, (r - (((r + 64) & 127) - 64)) >> 7)
        la, ha = sp(pad(ra))
        lb, hb = sp(pad(rb))
        acc0 = rng.integers(0, 1 << 14, (G, M, N)) % m
        acc = cu(acc0.astype(np.int16))
        R.accumulate(cu(ha.astype(np.int8)), cu(la.astype(np.int8)), cu(hb.astype(np.int8)),
                     cu(lb.astype(np.int8)), acc)
        ref = (acc0 + np.einsum("gik,gjk->gij", ra, rb)) % m
        _eq(host(acc).astype(np.int64), ref, f"product {M}x{N}x{K}")


def product_configs_case(dev_put, M=300, N=200, K=130, nb=2, L=10, configs=None, syrk=True):
    """every tile configuration of the product (the autotuner's choices) gives the same residues,
    equal to numpy int64; the syrk mode too on its lower tiles"""
    from sdpvibe_amd.rns import GEMM_CONFIGS, chunk_major
    R = mpmat.ctx(L)
    rng = np.random.default_rng(11)
    G = R.n
    m = np.array(R.mods, np.int64)[:, None, None, None]
    ra = rng.integers(-8128, 8129, (G, nb, M, K))
    rb = rng.integers(-8128, 8129, (G, nb, N, K))
    ra = np.where(np.abs(ra) <= (m - 1) // 2, ra, 0)
    rb = np.where(np.abs(rb) <= (m - 1) // 2, rb, 0)
    kp = R.kp(K)
    pad = lambda r: np.pad(r, ((0, 0), (0, 0), (0, 0), (0, kp - K))).reshape(G, -1, kp)
    sp = lambda r: (((r + 64) & 127) - 64, (r - (((r + 64) & 127) - 64)) >> 7)
    la, ha = sp(pad(ra))
    lb, hb = sp(pad(rb))
    A = [chunk_major(dev_put(x.astype(np.int8))) for x in (ha, la)]
    B = [chunk_major(dev_put(x.astype(np.int8))) for x in (hb, lb)]
    ref = np.einsum("gbik,gbjk->gbij", ra, rb) % m
    inv = np.array([int(v) for v in R.inv32.cpu()], np.int64)[:, None, None, None]
    yref = ref * inv % m
    yref = np.where(yref > m // 2, yref - m, yref)
    for cfg in GEMM_CONFIGS if configs is None else configs:
        acc = dev_put(np.zeros((G, nb * M, N), np.int16))
        R.product(A[0], A[1], B[0], B[1], acc, nb, M, N, nb * M * N, config=cfg)
        _eq(acc.cpu().numpy().astype(np.int64).reshape(G, nb, M, N), ref, f"config {cfg}")
        # CRT-ready output: centred residue * (M/m)^-1, then accumulated once more (2y)
        R.product(A[0], A[1], B[0], B[1], acc, nb, M, N, nb * M * N, config=cfg, scaled=True)
        _eq(acc.cpu().numpy().astype(np.int64).reshape(G, nb, M, N), yref, f"config {cfg} scaled")
        R.product(A[0], A[1], B[0], B[1], acc, nb, M, N, nb * M * N, config=cfg, scaled=True, accumulate=True)
        y2 = 2 * yref % m
        _eq(acc.cpu().numpy().astype(np.int64).reshape(G, nb, M, N), np.where(y2 > m // 2, y2 - m, y2),
            f"config {cfg} scaled accumulate")
    if not syrk:
        return
    # syrk: A = B, lower tiles, with the size rule (-1: 128 here, as the reference) and every square tile
    from sdpvibe_amd.rns import SYRK_CONFIGS
    A1 = [chunk_major(dev_put(x.astype(np.int8)[:, :M])) for x in (ha, la)]
    full = np.einsum("gik,gjk->gij", ra[:, 0], ra[:, 0]) % m[:, 0]
    low = np.tril(np.ones((M, M), bool))
    for cfg in (-1,) + SYRK_CONFIGS:
        acc = dev_put(np.zeros((G, M, M), np.int16))
        R.product(A1[0], A1[1], A1[0], A1[1], acc, 1, M, M, M * M, accumulate=True, syrk=True, config=cfg)
        got = acc.cpu().numpy().astype(np.int64)
        _eq(np.where(low, got, 0), np.where(low, full, 0), f"syrk lower, configuration {cfg}")
    with pytest.raises(ValueError):
        R.product(A1[0], A1[1], A1[0], A1[1], acc, 1, M, M, M * M, accumulate=True, syrk=True, config=12)


def carry_edge_case(dev_put, L):
    """exact products whose normalisation carries through every limb: 2(2^NB - 1) + 1 = 2^(NB+1) - 1
    rounds up to 2^(NB+1) (the rounding carry), and negative powers of two (the negation carry)"""
    NB = 32 * L - 2
    big = (1 << NB) - 1
    a = [[big, 1], [-big, -1], [-(1 << (NB - 1)), 0], [1 << (NB - 1), -(1 << (NB - 1))]]
    b = [[2, 1 << (NB - 1)], [1, 1 << (NB - 1)]]
    Am = to_planes([v for r in a for v in r], L).reshape(L, 1, 4, 2)
    Bm = to_planes([v for r in b for v in r], L).reshape(L, 1, 2, 2)
    Ae = np.array([[0, 0], [0, 0], [-5, EZ], [3, 3]], np.int64).reshape(1, 4, 2)
    Be = np.array([[0, 0], [0, 0]], np.int64).reshape(1, 2, 2)
    C = F.gemm(F.MPF(dev_put(Am), dev_put(Ae)), F.MPF(dev_put(Bm), dev_put(Be)))
    rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be)
    _eq(C.m.cpu().numpy(), rm, f"carry edge m L={L}")
    _eq(C.e.cpu().numpy(), re_, f"carry edge e L={L}")


@pytest.mark.parametrize("L", [1, 2, 3, 24, 25])
def test_gemm_carry_edges(L):
    carry_edge_case(cu, L)


def small_and_round_case(L):
    """products that normalise by a left shift (1, -1, -2^k), zero, and the fixed-exponent rounding
    (round half up of the two's complement value) at right, left and out-of-range shifts"""
    from tests.intutil import norm_closed_form, rnd_half_up, wrap
    NB = 32 * L - 2
    a = [[1, 0, 0], [-1, 0, 0], [0, 0, 0], [-(1 << (NB - 1)), 3, 0], [5, -7, 1 << (NB - 3)]]
    b = [[1, 0, 0], [1 << 5, 0, 0], [-3, 5, -1], [1 << (NB - 2), 1, -1]]
    exact = [[sum(x * y for x, y in zip(ra, rb)) for rb in b] for ra in a]
    A = mpmat.from_ints(a, 0, L)
    Bt = mpmat.from_ints(b, 0, L)
    # normalisation: value = m * 2^e with |m| of NB bits (zero: EZ)
    C = F.gemm(F.MPF(A.limbs.contiguous(), torch.zeros((5, 3), dtype=torch.int64, device=A.limbs.device)),
               F.MPF(Bt.limbs.permute(0, 2, 1).contiguous(),
                     torch.zeros((3, 4), dtype=torch.int64, device=A.limbs.device)))
    cm, ce = C.m.cpu().numpy(), C.e.cpu().numpy()
    got = from_planes(cm.reshape(L, -1))
    for idx, v in enumerate(v for row in exact for v in row):
        e = int(ce.reshape(-1)[idx])
        if v == 0:
            assert got[idx] == 0 and e == EZ
        else:
            assert (got[idx], e) == norm_closed_form(v, NB), (idx, v)
    for s in (0, 1, 31, 33, NB, 64 * L + 40, -1, -33):
        D = mpmat.gemm_nt(A, Bt, out_exp=s)
        dv = from_planes(D.limbs.cpu().numpy().reshape(L, -1))
        for idx, v in enumerate(v for row in exact for v in row):
            assert dv[idx] == wrap(rnd_half_up(v, s), L), (s, idx)


@pytest.mark.parametrize("L", [1, 2, 24])
def test_small_values_and_rounding(L):
    small_and_round_case(L)


def test_product_configs():
    product_configs_case(cu)


def _gemm_case(L, nb, M, K, N, transA, transB, seed):
    rng = random.Random(seed)
    shpA = (nb, K, M) if transA else (nb, M, K)
    shpB = (nb, N, K) if transB else (nb, K, N)
    a, ea = rand_mpf_ints(rng, int(np.prod(shpA)), L, 60)
    b, eb = rand_mpf_ints(rng, int(np.prod(shpB)), L, 60)
    rng.shuffle(a)
    Am = to_planes(a, L).reshape((L,) + shpA)
    Bm = to_planes(b, L).reshape((L,) + shpB)
    Ae = np.array(ea, np.int64).reshape(shpA)
    Be = np.array(eb, np.int64).reshape(shpB)
    Ae[np.asarray(a).reshape(shpA) == 0] = EZ
    Be[np.asarray(b).reshape(shpB) == 0] = EZ
    Ae.reshape(-1)[3] = EZ                            # an EZ entry with nonzero limbs: ignored
    if M > 2:
        Am[:, 0, 1] = 0                               # an all-zero row of op(A) when not transposed
        Ae[0, 1] = EZ
    return Am, Ae, Bm, Be


@pytest.mark.parametrize("L,nb,M,K,N", [(1, 1, 3, 4, 5), (2, 1, 33, 70, 31), (10, 2, 40, 33, 70),
                                         (16, 2, 20, 33, 40), (20, 1, 65, 40, 33), (24, 1, 130, 200, 129),
                                         (26, 3, 17, 64, 65), (28, 1, 33, 50, 64), (34, 1, 64, 65, 64),
                                         (2, 1, 5, 700, 9), (1, 1, 3, 5000, 4)])
@pytest.mark.parametrize("transA,transB", [(False, False), (True, False), (False, True), (True, True)])
def test_mpf_gemm(L, nb, M, K, N, transA, transB):
    Am, Ae, Bm, Be = _gemm_case(L, nb, M, K, N, transA, transB, L * 7919 + M + 2 * transA + transB)
    rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be, transA, transB)
    C = F.gemm(F.MPF(cu(Am), cu(Ae)), F.MPF(cu(Bm), cu(Be)), transA, transB)
    _eq(host(C.m), rm, "gemm m")
    _eq(host(C.e), re_, "gemm e")
    if nb == 1:
        C2 = F.gemm(F.MPF(cu(Am[:, 0]), cu(Ae[0])), F.MPF(cu(Bm[:, 0]), cu(Be[0])), transA, transB)
        _eq(host(C2.m), rm[:, 0], "gemm 2d m")


def test_mpf_gemm_512_L24():
    """The 512 x 512 x 512 product at 768 bits (L = 24) on Gaussian inputs, bit-exact."""
    L, n = 24, 512
    Am, Ae = gaussian_mpf((1, n, n), L, seed=0)
    Bm, Be = gaussian_mpf((1, n, n), L, seed=1)
    C = F.gemm(F.MPF(cu(Am), cu(Ae)), F.MPF(cu(Bm), cu(Be)))
    rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be)
    _eq(host(C.e), re_, "gemm512 e")
    _eq(host(C.m), rm, "gemm512 m")


@pytest.mark.parametrize("mode", ["col", "mat", -1000, "fixed"])
def test_gemm_nt_modes(mode):
    # tests/test_gemm.py:11-26 shapes at L=34
    L = 34
    rng = random.Random(1)
    for (M, N, K) in [(5, 7, 3), (40, 33, 70), (64, 64, 300), (130, 20, 1000)]:
        A = [[rng.getrandbits(32 * L - 40) * rng.choice([-1, 1]) for _ in range(K)] for _ in range(M)]
        B = [[rng.getrandbits(32 * L - 40) * rng.choice([-1, 1]) for _ in range(K)] for _ in range(N)]
        eA, eB = -1000, -1060
        Am, Bm = mpmat.from_ints(A, eA, L), mpmat.from_ints(B, eB, L)
        out_exp = (eA + eB + 1100) if mode == "fixed" else mode   # product fits 34 limbs
        C = mpmat.gemm_nt(Am, Bm, out_exp=out_exp)
        Ci, ce = mpmat.to_ints(C)
        W = O.gemm_exact(host(Am.limbs), host(Bm.limbs), (2 * 32 * L + 64) // 32 + 2)
        exact = [[v for v in from_planes(W.reshape(W.shape[0], -1))[i * N:(i + 1) * N]] for i in range(M)]
        for i in range(M):
            for j in range(N):
                assert exact[i][j] == sum(A[i][k] * B[j][k] for k in range(K))
                s = ce[j] - (eA + eB)
                ref = (exact[i][j] + (1 << (s - 1))) >> s if s > 0 else exact[i][j] << -s
                if mode == "mat":
                    assert abs(Ci[i][j] - ref) <= 1
                else:
                    assert Ci[i][j] == ref, (M, N, K, i, j)
