"""GPU parity of the batched layer (sdpvibe/batched.py) against the oracle."""
import random

import numpy as np
import pytest
import torch

from oracle import oracle as O
from sdpvibe_amd import batched as Bm
from tests.intutil import from_planes, to_planes

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}"


def _fx(rng, L, shape, bits):
    n = int(np.prod(shape))
    return to_planes([rng.randrange(-(1 << bits), 1 << bits) for _ in range(n)], L).reshape((L,) + shape)


@pytest.mark.parametrize("L,nb,M,N,K", [(3, 2, 4, 5, 6), (24, 3, 32, 32, 32), (24, 5, 40, 7, 96), (34, 2, 64, 100, 33),
                                         (24, 3, 40, 1, 300), (32, 2, 17, 2, 65), (5, 4, 1, 1, 1)])
@pytest.mark.parametrize("mode", ["fixed", "col", "mat", "rowexp_col", "rowexp_fixed"])
def test_bgemm_nt(L, nb, M, N, K, mode):
    rng = random.Random(L * 100 + M + N + K + len(mode))
    A = _fx(rng, L, (nb, M, K), 32 * L - 40)
    Bt = _fx(rng, L, (nb, N, K), 32 * L - 40)
    eA, eB = -(32 * L - 40), -(32 * L - 41)
    rowexp = np.array([[rng.randrange(-3000, -2000) for _ in range(N)] for _ in range(nb)], np.int64)
    use_row = mode.startswith("rowexp")
    e_in = (eA + rowexp) if use_row else np.full((nb, N), eA + eB, np.int64)
    Ab = Bm.BMP(cu(A), eA)
    Bb = Bm.BMP(cu(Bt), eB)
    if mode in ("fixed", "rowexp_fixed"):
        oe = int(e_in.max()) + 2 * (32 * L - 40) + 8 - (32 * L - 32)
        C = Bm.bgemm_nt(Ab, Bb, out_exp=oe, bt_rowexp=cu(rowexp) if use_row else None)
        ref, rexp = O.bgemm_nt(A, Bt, e_in, 0, np.full((nb, N), oe, np.int64))
        assert C.exp == oe
    else:
        omode = {"col": "col", "mat": "mat", "rowexp_col": "col"}[mode]
        C = Bm.bgemm_nt(Ab, Bb, out_exp=omode, bt_rowexp=cu(rowexp) if use_row else None)
        ref, rexp = O.bgemm_nt(A, Bt, e_in, 2 if omode == "mat" else 1)
        _eq(C.exp.cpu().numpy(), rexp, "bgemm exps")
    _eq(C.limbs.cpu().numpy(), ref, f"bgemm_nt {mode}")


@pytest.mark.parametrize("L", [2, 24])
def test_brescale_and_badd(L):
    rng = random.Random(L)
    nb, R, C = 3, 5, 7
    A = _fx(rng, L, (nb, R, C), 32 * L - 3)
    B = _fx(rng, L, (nb, R, C), 32 * L - 3)
    ea = np.array([[rng.randrange(-500, -400) for _ in range(C)] for _ in range(nb)], np.int64)
    eb = np.array([[rng.randrange(-500, -400) for _ in range(C)] for _ in range(nb)], np.int64)
    for sgn in (1, -1):
        out = Bm.badd(Bm.BMP(cu(A), cu(ea)), Bm.BMP(cu(B), cu(eb)), sgn, out_exp="align")
        ref, rEO = O.badd_align(A, B, ea, eb, sgn)
        _eq(out.limbs.cpu().numpy(), ref, "badd")
        _eq(out.exp.cpu().numpy(), rEO, "badd exp")
    same = Bm.badd(Bm.BMP(cu(A), -7), Bm.BMP(cu(B), -7), -1, out_exp=-7)
    _eq(same.limbs.cpu().numpy(), O.add_same(A.reshape(L, -1), B.reshape(L, -1), -1).reshape(A.shape), "badd same")
    ne = ea + 3
    r = Bm.brescale(Bm.BMP(cu(A), cu(ea)), cu(ne), L + 1)
    s = np.broadcast_to((ne - ea)[:, None, :], (nb, R, C)).reshape(-1)
    ref, _ = O.shift_round(A.reshape(L, -1), s, L + 1)
    _eq(r.limbs.cpu().numpy().reshape(L + 1, -1), ref, "brescale")


def _tile_inputs(L, nb, BT, logc, seed):
    from tests.intutil import int_to_words, spd_tile
    WL = L + 24
    tiles = [spd_tile(BT, logc, seed + b, 32 * L - 40) for b in range(nb)]
    flat = [tiles[b][i][j] for b in range(nb) for i in range(BT) for j in range(BT)]
    A = to_planes(flat, L).reshape(L, nb, BT, BT)
    W = np.zeros((nb, BT, BT, WL), np.uint32)
    for b in range(nb):
        for i in range(BT):
            for j in range(BT):
                W[b, i, j] = int_to_words(tiles[b][i][j] << 40, WL)
    return A, W


@pytest.mark.parametrize("L,nb,BT,logc", [(2, 2, 32, 6), (5, 1, 17, 10), (8, 3, 32, 30), (24, 4, 32, 60),
                                          (24, 2, 64, 40), (40, 1, 48, 90)])
def test_tile_chol(L, nb, BT, logc):
    """sdpv_tile_chol (widen, factor, invert, narrow) bit for bit against the oracle's
    chol_tiles/trinv_tiles restatement"""
    from sdpvibe_amd.basecase import tile_factor
    from tests.intutil import rnd_half_up, words_to_int, wrap
    A, W = _tile_inputs(L, nb, BT, logc, 7 * L + BT)
    Lw, _, X = O.tile_chol(W, L)
    Lo, Xw, _ = tile_factor(cu(A), BT)
    exp_L = to_planes([wrap(rnd_half_up(words_to_int(Lw[b, i, j]), 40), L) for b in range(nb) for i in range(BT)
                       for j in range(BT)], L).reshape(L, nb, BT, BT)
    _eq(Xw.cpu().numpy(), np.ascontiguousarray(X.transpose(3, 0, 1, 2)).view(np.int32), "tile inverse")
    _eq(Lo.cpu().numpy(), exp_L, "tile factor")


def test_bchol_tile_inverse_exponents():
    """bchol_tile: X rounded to one exponent per block, top of X at the headroom below 32L bits"""
    from sdpvibe_amd.mpmat import headroom_bits
    L, nb, BT = 24, 3, 32
    A, W = _tile_inputs(L, nb, BT, 40, 5)
    Lw, _, X = O.tile_chol(W, L)
    Lt, t = Bm.bchol_tile(Bm.BMP(cu(A), -(32 * L - 40)))
    ex = t.X.exp.cpu().numpy()
    assert (ex == ex[:, :1]).all()
    from tests.intutil import words_to_int
    for b in range(nb):
        xs = [words_to_int(X[b, i, j]) for i in range(BT) for j in range(BT)]
        top = max(abs(v).bit_length() if v >= 0 else (~v).bit_length() for v in xs) - 32 * L
        assert ex[b, 0] == top - headroom_bits(L)
        s = np.full(BT * BT, int(ex[b, 0]) + 32 * L, np.int64)
        ref, _ = O.shift_round(to_planes(xs, L + 24), s, L)
        _eq(t.X.limbs[:, b].reshape(L, -1).cpu().numpy(), ref, "bchol_tile X")


@pytest.mark.parametrize("L,nb,M,N,K", [(24, 1, 600, 1, 1500), (24, 60, 96, 1, 96), (32, 2, 100, 2, 3000),
                                         (63, 1, 7, 1, 64), (3, 300, 32, 1, 32)])
def test_narrow_products_equal_the_residue_path(L, nb, M, N, K, monkeypatch):
    """sdpv_exact_dot (one or two right-hand-side columns) gives bitwise the residue path's results,
    also from a strided view of the left operand (a sub-block of a larger factor) and at extreme
    limb values"""
    rng = random.Random(L + M + N + K)
    top = 1 << (32 * L - 1)
    vals = lambda n: [rng.choice([rng.randrange(-top, top), top - 1, -top, 0, -1]) for _ in range(n)]
    big = to_planes(vals(nb * (M + 3) * (K + 5)), L).reshape(L, nb, M + 3, K + 5)
    Bt = to_planes(vals(nb * N * K), L).reshape(L, nb, N, K)
    rowexp = cu(np.array([[rng.randrange(-900, 900) - 32 * L for _ in range(N)] for _ in range(nb)], np.int64))
    Av = Bm.BMP(cu(big)[:, :, 2:2 + M, 5:5 + K], -(32 * L - 8))            # strided view
    Bb = Bm.BMP(cu(Bt), 0)
    fixed = int(rowexp.max()) - (32 * L - 8) + 32 * L + 48                  # fits in L limbs
    for mode in ("col", fixed):
        monkeypatch.setattr(Bm, "NARROW_MK", 1 << 40)                      # the direct path at every size
        direct = Bm.bgemm_nt(Av, Bb, out_exp=mode, bt_rowexp=rowexp)
        monkeypatch.setattr(Bm, "NARROW", 0)                               # the residue path
        resid = Bm.bgemm_nt(Bm.BMP(Av.limbs.contiguous(), Av.exp), Bb, out_exp=mode, bt_rowexp=rowexp)
        monkeypatch.setattr(Bm, "NARROW", 2)
        _eq(direct.limbs.cpu().numpy(), resid.limbs.cpu().numpy(), f"narrow product limbs ({mode})")
        if mode == "col":
            _eq(direct.exp.cpu().numpy(), resid.exp.cpu().numpy(), "narrow product exponents")


def test_column_panels_and_k_chunks_equal_one_pass(monkeypatch):
    """mpf.gemm, mpmat.gemm_nt and bgemm_nt with the residue accumulator split into column panels (a
    small ACC_BUDGET) and the inner dimension into chunks (a small RES_BUDGET) give bitwise the
    one-pass results: the in-place panel reads (split offsets), the per-panel splits and the
    per-panel rounding are exercised"""
    from sdpvibe_amd import mpf as F
    from sdpvibe_amd import mpmat
    from sdpvibe_amd.synthetic import gaussian_mpf
    L, nb, M, K, N = 5, 2, 70, 300, 200
    am, ae = gaussian_mpf((nb, M, K), L, seed=31)
    bm, be = gaussian_mpf((nb, K, N), L, seed=32)
    A = F.MPF(torch.from_numpy(am).cuda(), torch.from_numpy(ae).cuda())
    B = F.MPF(torch.from_numpy(bm).cuda(), torch.from_numpy(be).cuda())
    rng = random.Random(33)
    Af = Bm.BMP(cu(_fx(rng, L, (nb, M, K), 32 * L - 40)), -(32 * L - 40))
    Bf = Bm.BMP(cu(_fx(rng, L, (nb, N, K), 32 * L - 40)), -(32 * L - 41))
    A2 = mpmat.MP(Af.limbs[:, 0].contiguous(), -1000)
    B2 = mpmat.MP(Bf.limbs[:, 1].contiguous(), -1060)
    one = F.gemm(A, B)
    oneb = Bm.bgemm_nt(Af, Bf, out_exp="col")
    onem = {mode: mpmat.gemm_nt(A2, B2, out_exp=mode) for mode in ("col", "mat", -1900)}
    R = F.ctx(L)
    monkeypatch.setattr(mpmat, "ACC_BUDGET", R.n * nb * M * 2 * 64)               # 64-column panels
    monkeypatch.setattr(mpmat, "RES_BUDGET", 2 * R.n * nb * (M + 64) * 2 * 64)    # 64-deep k chunks
    panels = F.gemm(A, B)
    panelsb = Bm.bgemm_nt(Af, Bf, out_exp="col")
    assert torch.equal(one.m, panels.m) and torch.equal(one.e, panels.e)
    assert torch.equal(oneb.limbs, panelsb.limbs) and torch.equal(oneb.exp, panelsb.exp)
    for mode, C in onem.items():
        D = mpmat.gemm_nt(A2, B2, out_exp=mode)
        assert torch.equal(C.limbs, D.limbs), mode
        if torch.is_tensor(C.exp):
            assert torch.equal(C.exp, D.exp), mode
        else:
            assert C.exp == D.exp, mode


@pytest.mark.parametrize("panels", [False, True])
def test_transposed_view_operand_equals_contiguous(panels, monkeypatch):
    """bgemm_nt with Bt given as the per-block transpose of a contiguous tensor (read through its
    strides, as the triangular solves pass B^T) gives the bits of the contiguous copy; with column
    panels forced, the panel offsets run through the strides too"""
    from sdpvibe_amd import mpmat
    L, nb, M, K, N = 6, 3, 40, 70, 150
    rng = random.Random(41)
    A = Bm.BMP(cu(_fx(rng, L, (nb, M, K), 32 * L - 40)), -(32 * L - 40))
    B = cu(_fx(rng, L, (nb, K, N), 32 * L - 40))                     # (L, nb, K, N): Bt = B^T per block
    rowexp = torch.from_numpy(np.array([[rng.randrange(-3000, -2000) for _ in range(N)] for _ in range(nb)],
                                       np.int64)).cuda()
    if panels:
        R = mpmat.ctx(L)
        monkeypatch.setattr(mpmat, "ACC_BUDGET", R.n * nb * M * 2 * 64)
    for mode in ("col", "mat", -(64 * L)):
        view = Bm.bgemm_nt(A, Bm.BMP(B.transpose(2, 3), 0), out_exp=mode, bt_rowexp=rowexp)
        copy = Bm.bgemm_nt(A, Bm.BMP(B.transpose(2, 3).contiguous(), 0), out_exp=mode, bt_rowexp=rowexp)
        assert torch.equal(view.limbs, copy.limbs), mode
        if torch.is_tensor(copy.exp):
            assert torch.equal(view.exp, copy.exp), mode
