"""Runs sdpvibe_amd operations through the emulated kernels (ASan) and compares them with the
oracle. Usage: python tests/emu/run_cases.py <case> [<case> ...]   (run with asan_env())."""
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from tests.emu import harness  # noqa: E402

harness.install()
from oracle import oracle as O  # noqa: E402
from sdpvibe_amd import mpf as F  # noqa: E402
from sdpvibe_amd import mpmat  # noqa: E402
from tests.intutil import EZ, rand_mpf_ints, to_planes  # noqa: E402


def gemm_case(L, nb, M, K, N, transA, transB, seed=0):
    rng = random.Random(seed)
    shpA = (nb, K, M) if transA else (nb, M, K)
    shpB = (nb, N, K) if transB else (nb, K, N)
    a, ea = rand_mpf_ints(rng, int(np.prod(shpA)), L, 60)
    b, eb = rand_mpf_ints(rng, int(np.prod(shpB)), L, 60)
    Am = to_planes(a, L).reshape((L,) + shpA)
    Bm = to_planes(b, L).reshape((L,) + shpB)
    Ae = np.array(ea, np.int64).reshape(shpA)
    Be = np.array(eb, np.int64).reshape(shpB)
    Ae[np.asarray(a).reshape(shpA) == 0] = EZ
    Be[np.asarray(b).reshape(shpB) == 0] = EZ
    C = F.gemm(F.MPF(torch.from_numpy(Am), torch.from_numpy(Ae)), F.MPF(torch.from_numpy(Bm), torch.from_numpy(Be)),
               transA, transB)
    rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be, transA, transB)
    ok = np.array_equal(C.m.numpy(), rm) and np.array_equal(C.e.numpy(), re_)
    print(f"gemm L={L} nb={nb} {M}x{K}x{N} tA={transA} tB={transB}: {'ok' if ok else 'MISMATCH'}", flush=True)
    return ok


CASES = {
    "gemm_small": lambda: gemm_case(1, 1, 3, 4, 5, False, False),
    "gemm_trans": lambda: all(gemm_case(2, 1, 5, 7, 6, ta, tb, 3) for ta in (False, True) for tb in (False, True)),
    "gemm_batched": lambda: gemm_case(3, 2, 33, 20, 40, False, True, 4),
    # the CRT's register-held words: 5, 6 and 8 digit blocks (7: the L = 24 cases)
    "gemm_regwords": lambda: all(gemm_case(L, 1, 9, 12, 10, False, True, L) for L in (16, 20, 28)),
    # column row exponents of B: 16-row blocks (few rows, k <= 4096), and k split over blocks with
    # a filled Erow and an atomic maximum (few rows, k > 4096)
    "gemm_ksplit": lambda: gemm_case(2, 1, 3, 600, 5, False, False, 9),
    "gemm_ksplit_long": lambda: gemm_case(1, 1, 2, 5000, 3, False, False, 13),
}

def gpu_test_gemm_cases():
    """the exact inputs of tests/test_gpu_kernels.py::test_mpf_gemm"""
    from tests.test_gpu_kernels import _gemm_case
    ok = True
    for (L, nb, M, K, N) in [(1, 1, 3, 4, 5), (2, 1, 33, 70, 31), (10, 2, 40, 33, 70)]:
        for ta in (False, True):
            for tb in (False, True):
                Am, Ae, Bm, Be = _gemm_case(L, nb, M, K, N, ta, tb, L * 7919 + M + 2 * ta + tb)
                rm, re_ = O.mpf_gemm(Am, Ae, Bm, Be, ta, tb)
                C = F.gemm(F.MPF(torch.from_numpy(Am), torch.from_numpy(Ae)),
                           F.MPF(torch.from_numpy(Bm), torch.from_numpy(Be)), ta, tb)
                good = np.array_equal(C.m.numpy(), rm) and np.array_equal(C.e.numpy(), re_)
                print(f"gpu-test gemm L={L} nb={nb} {M}x{K}x{N} tA={ta} tB={tb}: {'ok' if good else 'MISMATCH'}",
                      flush=True)
                ok &= good
    return ok


CASES["gpu_gemm"] = gpu_test_gemm_cases


def bgemm_cases():
    from sdpvibe_amd import batched as Bm
    from tests.intutil import to_planes as tp
    ok = True
    L, nb, M, N, K = 3, 2, 33, 7, 20
    rng = random.Random(5)
    A = tp([rng.randrange(-(1 << 56), 1 << 56) for _ in range(nb * M * K)], L).reshape(L, nb, M, K)
    Bt = tp([rng.randrange(-(1 << 56), 1 << 56) for _ in range(nb * N * K)], L).reshape(L, nb, N, K)
    eA, eB = -56, -57
    e_in = np.full((nb, N), eA + eB, np.int64)
    for mode, om in (("col", 1), ("mat", 2), (40, 0)):
        C = Bm.bgemm_nt(Bm.BMP(torch.from_numpy(A), eA), Bm.BMP(torch.from_numpy(Bt), eB), out_exp=mode)
        ref, rexp = O.bgemm_nt(A, Bt, e_in, om, np.full((nb, N), 40, np.int64) if om == 0 else None)
        good = np.array_equal(C.limbs.numpy(), ref)
        if om != 0:
            good &= np.array_equal(C.exp.numpy(), rexp)
        print(f"bgemm_nt mode={mode}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok &= good
    return ok


CASES["bgemm"] = bgemm_cases


def bgemm_narrow_cases():
    """bgemm_nt with one or two right-hand-side columns (the direct exact products) against the oracle:
    every output mode, per-column input exponents, k over several chunks, extreme limb values"""
    from sdpvibe_amd import batched as Bm
    from tests.intutil import to_planes as tp
    ok = True
    for L, nb, M, N, K, seed in ((3, 2, 9, 1, 70, 1), (3, 1, 5, 2, 600, 2), (24, 1, 4, 1, 66, 3), (2, 3, 1, 2, 1, 4)):
        rng = random.Random(seed)
        top = 1 << (32 * L - 1)
        pick = lambda: rng.choice([rng.randrange(-top, top), top - 1, -top, 0, -1, 1 << rng.randrange(32 * L - 1)])
        A = tp([pick() for _ in range(nb * M * K)], L).reshape(L, nb, M, K)
        Bt = tp([pick() for _ in range(nb * N * K)], L).reshape(L, nb, N, K)
        eA = -(32 * L - 8)
        eB = np.array([[rng.randrange(-40, 40) - 32 * L for _ in range(N)] for _ in range(nb)], np.int64)
        e_in = eA + eB
        for mode, om in (("col", 1), ("mat", 2), (17, 0)):
            C = Bm.bgemm_nt(Bm.BMP(torch.from_numpy(A), eA), Bm.BMP(torch.from_numpy(Bt), 0), out_exp=mode,
                            bt_rowexp=torch.from_numpy(eB))
            ref, rexp = O.bgemm_nt(A, Bt, e_in, om, np.full((nb, N), 17, np.int64) if om == 0 else None)
            good = np.array_equal(C.limbs.numpy(), ref)
            if om != 0:
                good &= np.array_equal(C.exp.numpy(), rexp)
            print(f"bgemm_nt narrow L={L} nb={nb} M={M} N={N} K={K} mode={mode}: {'ok' if good else 'MISMATCH'}",
                  flush=True)
            ok &= good
    return ok


CASES["bgemm_narrow"] = bgemm_narrow_cases


def tile_case(L, nb, BT, logc, seed):
    """sdpv_tile_chol against the oracle's chol_tiles/trinv_tiles restatement, bit for bit"""
    from sdpvibe_amd.basecase import tile_factor
    from tests.intutil import int_to_words, rnd_half_up, spd_tile, words_to_int, wrap
    WL = L + 24
    tiles = [spd_tile(BT, logc, seed + b, 32 * L - 40) for b in range(nb)]
    flat = [tiles[b][i][j] for b in range(nb) for i in range(BT) for j in range(BT)]
    A = to_planes(flat, L).reshape(L, nb, BT, BT)
    W = np.zeros((nb, BT, BT, WL), np.uint32)
    for b in range(nb):
        for i in range(BT):
            for j in range(BT):
                W[b, i, j] = int_to_words(tiles[b][i][j] << 40, WL)
    Lw, _, X = O.tile_chol(W, L)
    Lo, Xw, _ = tile_factor(torch.from_numpy(A), BT)
    exp_L = to_planes([wrap(rnd_half_up(words_to_int(Lw[b, i, j]), 40), L) for b in range(nb) for i in range(BT)
                       for j in range(BT)], L).reshape(L, nb, BT, BT)
    exp_X = np.ascontiguousarray(X.transpose(3, 0, 1, 2)).view(np.int32)
    okL = np.array_equal(Lo.numpy(), exp_L)
    okX = np.array_equal(Xw.numpy(), exp_X)
    print(f"tile_chol L={L} nb={nb} BT={BT} logc={logc}: L {'ok' if okL else 'MISMATCH'}, "
          f"X {'ok' if okX else 'MISMATCH'}", flush=True)
    if not okX:
        bad = np.argwhere(Xw.numpy() != exp_X)
        print("first X mismatches (word, b, i, j):", bad[:8].tolist(), flush=True)
    return okL and okX


CASES["tile"] = lambda: tile_case(2, 2, 32, 6, 11) and tile_case(4, 1, 20, 12, 3)
# the lane-parallel pivot path up to its limit (WL = L + 24 = 64) and the wave path just beyond it
CASES["tile_wide"] = lambda: (tile_case(24, 1, 32, 20, 5) and tile_case(40, 1, 12, 8, 7)
                              and tile_case(41, 1, 8, 4, 9))


def chol_cases():
    from tests.test_gpu_chol import chol_parity
    import tests.test_gpu_chol as T
    T.cu = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    chol_parity(2, 2, 40, 3, 5, seed=1)
    print("chol: ok", flush=True)
    return True


CASES["chol"] = chol_cases


def crt_cases():
    """RNS.crt_block at L=1, 24 and 130 (n > 512 primes: fewer waves per workgroup)"""
    from sdpvibe_amd import mpmat as MM
    from tests.intutil import from_planes as fp
    ok = True
    for L in (1, 24, 130):
        R = MM.ctx(L)
        rng = random.Random(41 + L)
        rows, N, start, width = 5, 13, 2, 11
        limit = R.M // 4
        vals = [[rng.randrange(-limit + 1, limit) for _ in range(N)] for _ in range(rows)]
        vals[0][2:7] = [0, 1, -1, limit - 1, -limit + 1]
        acc = np.array([[[v % p for v in row] for row in vals] for p in R.mods], np.int16)
        W, bl, Bq = R.crt_block(torch.from_numpy(acc), N, start, width, rows)
        exp = [v for row in vals for v in row[start:start + width]]
        sel = acc[:, :, start:start + width].reshape(R.n, -1)
        rW, rbl = O.crt(sel, np.array(R.mods), R.E // 4)
        good = np.array_equal(W.numpy(), rW) and np.array_equal(bl.numpy(), rbl) and fp(W.numpy()) == exp
        from tests.test_gpu_kernels import scaled_residues
        W2, bl2, _ = R.crt_block(torch.from_numpy(scaled_residues(R, acc)), N, start, width, rows, scaled=True)
        good &= np.array_equal(W2.numpy(), rW) and np.array_equal(bl2.numpy(), rbl)
        print(f"crt_block L={L} n={R.n}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok &= good
    return ok


CASES["crt"] = crt_cases


def residue_cases():
    """RNS.residues at L = 3 and L = 100 (no limb staging in LDS) against the oracle"""
    from sdpvibe_amd import mpmat as MM
    ok = True
    for L, K in ((3, 17), (100, 20)):
        R = MM.ctx(L)
        rng = random.Random(L)
        rows = 5
        X = to_planes([rng.randrange(-(1 << (32 * L - 1)), 1 << (32 * L - 1)) for _ in range(rows * K)], L)
        X = X.reshape(L, rows, K)
        hi, lo = R.residues(torch.from_numpy(X), 3, K - 5)
        got = hi.numpy().astype(np.int64) * 128 + lo.numpy().astype(np.int64)
        ref = O.residues(X, 3, K - 5, np.array(R.mods))
        good = np.array_equal(got[:, :, :K - 5], ref) and not got[:, :, K - 5:].any()
        print(f"residues L={L}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok &= good
    return ok


CASES["residues"] = residue_cases


def gram_cases():
    import tests.test_gpu_gram as T
    T.cu = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    T.gram_parity(3, 40, (5, 70), [0, 0], seed=1)
    assert T.gram_parity(2, 33, (20, 9, 11), [0, 60, -5], seed=2) == 1
    T.gram_parity(2, 20, (10, 10), [0, 3], seed=3, e_ref=True)
    print("gram: ok", flush=True)
    return True


CASES["gram"] = gram_cases


def linalg_cases():
    import tests.test_gpu_linalg as T
    T.cu = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    T.linalg_chol_inv_case(2, 40, 4)
    T.linalg_chol_case(2, 64, "subst", 0, m=2)
    T.test_trsv_tile(2, 3)
    print("linalg: ok", flush=True)
    return True


CASES["linalg"] = linalg_cases


def eig_cases():
    import tests.test_gpu_steplen as T
    from sdpvibe_amd import solver as S
    orig = torch.Tensor.cuda
    torch.Tensor.cuda = lambda self, *a, **k: self
    try:
        for n in (1, 2, 7, 40):
            T.test_symeig_matches_numpy(n)
        T.test_mpf_from_float_exact()
    finally:
        torch.Tensor.cuda = orig
    A = F.from_mp([mp_val for mp_val in ["1.5", "-2.25e300", "3e-320", "0", "7", "-1e-5"]], 24, (1, 2, 3))
    d = S.to_double(A, torch.tensor([0], dtype=torch.int64))
    assert np.array_equal(d.numpy(), A.to_float())
    print("eig: ok", flush=True)
    return True


CASES["eig"] = eig_cases


def product_cases(configs, syrk):
    from tests.test_gpu_kernels import product_configs_case
    product_configs_case(lambda a: torch.from_numpy(np.ascontiguousarray(a)), M=70, N=40, K=65, nb=2, L=2,
                         configs=configs, syrk=syrk)
    print(f"product configs {configs}: ok", flush=True)
    return True


# the twelve tile configurations in two halves, so that the halves can run side by side
CASES["product_a"] = lambda: product_cases(range(0, 6), True)
CASES["product_b"] = lambda: product_cases(range(6, 13), False)


def product_staged_cases():
    """the configuration with the LDS-staged epilogue (13) at a size with full tiles (its staged
    path) and edge tiles (the per-entry path)"""
    from tests.test_gpu_kernels import product_configs_case
    put = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    product_configs_case(put, M=130, N=136, K=65, nb=1, L=1, configs=[13], syrk=False)
    # N not a multiple of 8: full tiles store entry by entry as well
    product_configs_case(put, M=70, N=133, K=64, nb=2, L=1, configs=[13], syrk=False)
    print("product config 13: ok", flush=True)
    return True


CASES["product_c"] = product_staged_cases


def carry_cases():
    from tests.test_gpu_kernels import carry_edge_case
    for L in (1, 2, 3, 24, 25):
        carry_edge_case(lambda a: torch.from_numpy(np.ascontiguousarray(a)), L)
    print("carry edges: ok", flush=True)
    return True


CASES["carry"] = carry_cases


def small_round_cases():
    from tests.test_gpu_kernels import small_and_round_case
    for L in (1, 2, 24):
        small_and_round_case(L)
    print("small values and rounding: ok", flush=True)
    return True


CASES["small_round"] = small_round_cases


def panel_cases():
    """mpf.gemm, mpmat.gemm_nt and bgemm_nt with column panels (in-place split offsets) and k chunks
    forced by small budgets: bitwise the one-pass results, and mpf.gemm equal to the oracle"""
    from sdpvibe_amd import batched as Bm
    L, nb, M, K, N = 2, 2, 9, 150, 140
    rng = random.Random(21)
    a, ea = rand_mpf_ints(rng, nb * M * K, L, 60)
    b, eb = rand_mpf_ints(rng, nb * K * N, L, 60)
    Am, Bmm = to_planes(a, L).reshape(L, nb, M, K), to_planes(b, L).reshape(L, nb, K, N)
    Ae, Be = np.array(ea, np.int64).reshape(nb, M, K), np.array(eb, np.int64).reshape(nb, K, N)
    Ae[np.asarray(a).reshape(nb, M, K) == 0] = EZ
    Be[np.asarray(b).reshape(nb, K, N) == 0] = EZ
    A = F.MPF(torch.from_numpy(Am), torch.from_numpy(Ae))
    B = F.MPF(torch.from_numpy(Bmm), torch.from_numpy(Be))
    fx = lambda shape: torch.from_numpy(np.ascontiguousarray(to_planes(
        [rng.randrange(-(1 << (32 * L - 40)), 1 << (32 * L - 40)) for _ in range(int(np.prod(shape)))], L)
        .reshape((L,) + shape)))
    Af, Bf = Bm.BMP(fx((nb, M, K)), -(32 * L - 40)), Bm.BMP(fx((nb, N, K)), -(32 * L - 41))
    A2, B2 = mpmat.MP(Af.limbs[:, 0].contiguous(), -100), mpmat.MP(Bf.limbs[:, 1].contiguous(), -106)
    one, oneb, onem = F.gemm(A, B), Bm.bgemm_nt(Af, Bf, out_exp="col"), mpmat.gemm_nt(A2, B2, out_exp="col")
    R = F.ctx(L)
    budgets = mpmat.ACC_BUDGET, mpmat.RES_BUDGET
    mpmat.ACC_BUDGET, mpmat.RES_BUDGET = R.n * nb * M * 2 * 64, 2 * R.n * nb * (M + 64) * 2 * 64
    try:
        pan, panb, panm = F.gemm(A, B), Bm.bgemm_nt(Af, Bf, out_exp="col"), mpmat.gemm_nt(A2, B2, out_exp="col")
    finally:
        mpmat.ACC_BUDGET, mpmat.RES_BUDGET = budgets
    rm, re_ = O.mpf_gemm(Am, Ae, Bmm, Be, False, False)
    ok = (torch.equal(one.m, pan.m) and torch.equal(one.e, pan.e) and np.array_equal(pan.m.numpy(), rm)
          and np.array_equal(pan.e.numpy(), re_) and torch.equal(oneb.limbs, panb.limbs)
          and torch.equal(oneb.exp, panb.exp) and torch.equal(onem.limbs, panm.limbs)
          and torch.equal(onem.exp, panm.exp))
    print(f"column panels and k chunks: {'ok' if ok else 'MISMATCH'}", flush=True)
    return ok


CASES["panels"] = panel_cases


def ksum_cases():
    """mpf.KSum over row segments against one gemm over the stacked segments, bitwise"""
    from tests.test_gpu_stream import ksum_case
    put = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    ksum_case(put, 3, 40, 1, [7, 0, 30, 1], seed=5)
    ksum_case(put, 2, 9, 3, [70, 65], seed=8)
    print("ksum: ok", flush=True)
    return True


CASES["ksum"] = ksum_cases


def bucketed_cases():
    from tests.test_gpu_bucketed import bucketed_case
    bucketed_case(lambda a: torch.from_numpy(np.ascontiguousarray(a)), 2, [3, 33, 9], 2, seed=5)
    print("bucketed chol: ok", flush=True)
    return True


CASES["bucketed"] = bucketed_cases


if __name__ == "__main__":  # noqa: E305
    ok = True
    for c in sys.argv[1:] or list(CASES):
        ok &= bool(CASES[c]())
    sys.exit(0 if ok else 1)

