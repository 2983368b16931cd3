"""Pins the C oracle (oracle/liboracle.so) on the CPU against the reference's own known-answer
tests, restated as Python-integer closed forms, and against mathematics (exact products,
error bounds in bits). No GPU needed."""
import random

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O
from tests.intutil import (EZ, MASK, bitlen_tc, from_planes, norm_closed_form, rand_mpf_ints, rnd_half_up,
                           to_planes, wrap)


def test_bitlen_at_powers_of_two():
    # tests/test_rns_kernels.py:16-25
    values = [v for bit in range(1, 95) for v in
              ((1 << bit) - 1, 1 << bit, (1 << bit) + 1, -(1 << bit) + 1, -(1 << bit), -(1 << bit) - 1)]
    values += [0, 1, -1]
    got = O.bitlen(to_planes(values, 3)).tolist()
    assert got == [bitlen_tc(v) for v in values]


def _flag_closed(y, Lout):
    w = [(y >> (32 * k)) & MASK for k in range(Lout + 1)]
    want = MASK if w[Lout - 1] >> 31 else 0
    return int(w[Lout] != want)


@pytest.mark.parametrize("Lin,Lout", [(1, 1), (3, 2), (5, 7), (24, 24), (24, 25)])
def test_shift_round_closed_form(Lin, Lout):
    rng = random.Random(5 + Lin * 31 + Lout)
    lim = 1 << (32 * Lin - 1)
    vals = [0, 1, -1, lim - 1, -lim, 2, -2, 3, -3] + [rng.randrange(-lim, lim) for _ in range(300)]
    shifts = [0, 1, -1, 31, 32, 33, -31, -32, -33, 32 * Lin - 1, 32 * Lin, 32 * Lin + 5, 1 << 30,
              -(32 * Lout), -(32 * Lout + 40), -(1 << 20)]
    shifts += [rng.randrange(-40 * Lout, 40 * Lin) for _ in range(len(vals) - len(shifts))]
    X = to_planes(vals, Lin)
    out, flag = O.shift_round(X, np.array(shifts, np.int64), Lout)
    got = from_planes(out)
    for v, s, g, f in zip(vals, shifts, got, flag.tolist()):
        y = rnd_half_up(v, s)
        assert g == wrap(y, Lout), (v, s)
        assert f == _flag_closed(y, Lout), (v, s)


@pytest.mark.parametrize("L", [1, 4, 24])
@pytest.mark.parametrize("sgn", [1, -1])
def test_add_same(L, sgn):
    rng = random.Random(L)
    lim = 1 << (32 * L - 1)
    a = [rng.randrange(-lim, lim) for _ in range(100)] + [lim - 1, -lim, 0]
    b = [rng.randrange(-lim, lim) for _ in range(100)] + [lim - 1, lim - 1, -lim]
    got = from_planes(O.add_same(to_planes(a, L), to_planes(b, L), sgn))
    assert got == [wrap(x + sgn * y, L) for x, y in zip(a, b)]


@pytest.mark.parametrize("L", [1, 4, 28, 34])
@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("uniform", [False, True])
def test_badd_align_closed_form(L, sign, uniform):
    # tests/test_badd.py:10-49
    rng = random.Random(19)
    nb, rows, cols = 2, 3, 5
    limit = 1 << (32 * L - 1)
    edge = [0, 1, -1, limit - 1, -limit]
    a = [rng.randrange(-limit, limit) for _ in range(nb * rows * cols)]
    b = list(reversed(a))
    a[:5], b[:5] = edge, list(reversed(edge))
    a[10:14] = [limit - 1, -limit, limit - 1, -limit]
    b[10:14] = [-limit, limit - 1, limit - 1, -limit]
    ea = [-100] * (nb * cols)
    eb = [-100] * (nb * cols) if uniform else [-100 + d for d in (0, 1, 31, 32, 33, 63, 64, 65, -10000, 10000)]
    A = to_planes(a, L).reshape(L, nb, rows, cols)
    B = to_planes(b, L).reshape(L, nb, rows, cols)
    out, EO = O.badd_align(A, B, np.array(ea).reshape(nb, cols), np.array(eb).reshape(nb, cols), sign)
    got = from_planes(out.reshape(L, -1))
    for i, (av, bv) in enumerate(zip(a, b)):
        j = (i // (rows * cols)) * cols + i % cols
        ex = max(ea[j], eb[j]) + 1
        expected = rnd_half_up(av, ex - ea[j]) + sign * rnd_half_up(bv, ex - eb[j])
        assert got[i] == wrap(expected, L)
        assert EO.reshape(-1)[j] == ex


@pytest.mark.parametrize("L,LI", [(1, 3), (2, 2), (10, 24), (24, 54), (26, 160)])
def test_mpf_norm_closed_form(L, LI):
    rng = random.Random(L * 7 + LI)
    lim = 1 << (32 * LI - 1)
    vals = [0, 1, -1, lim - 1, -lim, 1 << (32 * L - 3), -(1 << (32 * L))]
    vals += [rng.randrange(-lim, lim) >> rng.randrange(0, 32 * LI) for _ in range(200)]
    vals = [v for v in vals if -lim <= v < lim]
    ein = [rng.randrange(-5000, 5000) for _ in vals]
    Om, eO = O.mpf_norm(to_planes(vals, LI), np.array(ein, np.int64), L)
    got = from_planes(Om)
    NB = 32 * L - 2
    for v, e0, g, ge in zip(vals, ein, got, eO.tolist()):
        r, sh = norm_closed_form(v, NB)
        if sh is None:
            assert g == 0 and ge == EZ
        else:
            assert g == wrap(r, L) and ge == e0 + sh


@pytest.mark.parametrize("L", [1, 2, 10, 26, 28, 34, 64])
@pytest.mark.parametrize("scalar", [False, True])
def test_mpf_mul_closed_form(L, scalar):
    # tests/test_mpf_mul.py:15-39
    rng = random.Random(71)
    bits = 32 * L - 2
    edge = [0, 1, -1, (1 << bits) - 1, -((1 << bits) - 1), 1 << (bits - 1), -(1 << (bits - 1)),
            (1 << (bits - 1)) + 1]
    values = edge + [rng.randrange(-(1 << bits), 1 << bits) for _ in range(129)]
    ea = [rng.randrange(-2000, 2000) for _ in values]
    bvals = [(1 << bits) - 1] if scalar else list(reversed(values))
    eb = [rng.randrange(-2000, 2000) for _ in bvals]
    Om, eO = O.mpf_mul(to_planes(values, L), np.array(ea), to_planes(bvals, L), np.array(eb))
    got = from_planes(Om)
    for i, actual in enumerate(got):
        j = 0 if scalar else i
        product = values[i] * bvals[j]
        r, sh = norm_closed_form(product, bits)
        if sh is None:
            assert actual == 0 and eO[i] == EZ
            continue
        assert actual == r
        assert eO[i] == ea[i] + eb[j] + sh


def _mpf_add_closed(a, ea, b, eb, sgn, L):
    """sdpvibe/mpf.py:92-124 restated on Python integers"""
    NB = 32 * L - 2
    W = L + 2
    if a == 0:
        ea = EZ
    if b == 0:
        eb = EZ
    if ea == EZ and eb == EZ:
        return 0, EZ
    na, nbg = a < 0, (b < 0) ^ (sgn < 0)
    t = max(ea, eb) - 32
    cl = 32 * (L + 3)

    def sm(mag, e):
        d = e - t
        s = cl if d < -cl else -d
        r = rnd_half_up(mag, s)          # magnitude: half up == half toward +inf
        return r % (1 << (32 * W))

    wa = 0 if ea == EZ else sm(abs(a), ea)
    wb = 0 if eb == EZ else sm(abs(b), eb)
    v = wrap((-wa if na else wa) + (-wb if nbg else wb), W)
    r, sh = norm_closed_form(v, NB)
    if sh is None:
        return 0, EZ
    return wrap(r, L), t + sh


@pytest.mark.parametrize("L", [1, 2, 10, 24, 34])
@pytest.mark.parametrize("sgn", [1, -1])
def test_mpf_add_closed_form_and_error(L, sgn):
    rng = random.Random(100 + L)
    n = 200
    a, ea = rand_mpf_ints(rng, n, L, 100)
    b, eb = rand_mpf_ints(rng, n, L, 100)
    b[20], eb[20] = -sgn * a[20], ea[20]                 # exact cancellation
    b[21], eb[21] = a[21], ea[21] + 5000                 # far apart
    Om, eO = O.mpf_add(to_planes(a, L), np.array(ea), to_planes(b, L), np.array(eb), sgn)
    got = from_planes(Om)
    mp.mp.prec = 4 * 32 * L + 200
    NB = 32 * L - 2
    for i in range(n):
        r, e = _mpf_add_closed(a[i], ea[i], b[i], eb[i], sgn, L)
        assert (got[i], int(eO[i])) == (r, e), i
        exact = mp.mpf(a[i]) * mp.mpf(2) ** ea[i] + sgn * mp.mpf(b[i]) * mp.mpf(2) ** eb[i]
        val = mp.mpf(got[i]) * mp.mpf(2) ** int(eO[i]) if got[i] else mp.mpf(0)
        # two roundings (alignment to L+2 words, then to NB bits): |err| <= 2^(2-NB) |exact|
        # plus the alignment residue 2^(t-1) of the smaller operand
        if exact != 0:
            t = max(ea[i] if a[i] else EZ, eb[i] if b[i] else EZ) - 32
            assert abs(val - exact) <= mp.mpf(2) ** (2 - NB) * abs(exact) + mp.mpf(2) ** (t - 1), i


def test_residues_against_integers():
    # tests/test_rns_kernels.py:28-52 (any list of primes: the oracle is generic in the moduli)
    from sdpvibe_amd.rns import PMAX, PMIN, _primes_desc
    mods = _primes_desc(PMAX, PMIN)[:40]
    for L in (1, 10, 28, 34):
        for K in (17, 65):
            rng = random.Random(123 + L + K)
            lim = 1 << (32 * L - 1)
            values = [[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(3)]
            values[0][3:8] = [0, 1, -1, -lim, lim - 1]
            X = np.stack([to_planes(r, L) for r in values], axis=1)          # (L, 3, K)
            got = O.residues(X, 3, K - 5, np.array(mods))
            for gi, m in enumerate(mods):
                for i, row in enumerate(values):
                    exp = [v % m for v in row[3:K - 2]]
                    exp = [r - m if r > m // 2 else r for r in exp]
                    assert got[gi, i].tolist() == exp


@pytest.mark.parametrize("nmods", [1, 7, 40, 112])
def test_crt_against_integers(nmods):
    # tests/test_rns_kernels.py:55-80
    from sdpvibe_amd.rns import PMAX, PMIN, _primes_desc
    mods = _primes_desc(PMAX, PMIN)[:nmods]
    M = 1
    for m in mods:
        M *= m
    rng = random.Random(41)
    limit = M // 4
    vals = [rng.randrange(-limit + 1, limit) for _ in range(60)]
    vals[:5] = [0, 1, -1, limit - 1, -limit + 1]
    vals[5:10] = [1 << b for b in (0, 31, 32, 63, 64)] if M > 1 << 70 else [2, 3, 4, 5, 6]
    vals[10:15] = [-(1 << b) for b in (0, 31, 32, 63, 64)] if M > 1 << 70 else [-2, -3, -4, -5, -6]
    acc = np.array([[v % m for v in vals] for m in mods], np.int16)
    nw = (M.bit_length() + 8) // 32 + 1
    W, bl = O.crt(acc, np.array(mods), nw)
    got = from_planes(W)
    assert got == vals
    assert bl.tolist() == [bitlen_tc(v) for v in vals]


def _py_mpf_gemm(Ai, Ae, Bi, Be, L):
    """mpf.gemm (sdpvibe/mpf.py:390-452) on Python ints: A (M,K), B (K,N) lists of (int, exp)."""
    NB = 32 * L - 2
    M, K, N = len(Ai), len(Ai[0]), len(Bi[0])

    def top(v, e):
        return EZ if e == EZ else e + bitlen_tc(v)

    def align(vals, exps):
        E = max(top(v, e) for v, e in zip(vals, exps)) - NB
        if E < EZ // 2:
            E = 0
        out = [wrap(rnd_half_up(v, (1 << 30) if e == EZ else E - e), L) for v, e in zip(vals, exps)]
        return out, E
    ra = [align(Ai[i], Ae[i]) for i in range(M)]
    cb = [align([Bi[k][j] for k in range(K)], [Be[k][j] for k in range(K)]) for j in range(N)]
    C = []
    for i in range(M):
        row = []
        for j in range(N):
            v = sum(x * y for x, y in zip(ra[i][0], cb[j][0]))
            r, sh = norm_closed_form(v, NB)
            row.append((0, EZ) if sh is None else (wrap(r, L), ra[i][1] + cb[j][1] + sh))
        C.append(row)
    return C


@pytest.mark.parametrize("L,M,K,N", [(1, 3, 4, 5), (3, 7, 9, 4), (10, 5, 12, 6)])
@pytest.mark.parametrize("transA,transB", [(False, False), (True, False), (False, True), (True, True)])
def test_mpf_gemm_oracle_against_python(L, M, K, N, transA, transB):
    rng = random.Random(L * 1000 + M * 100 + K * 10 + N + 2 * transA + transB)
    a, ea = rand_mpf_ints(rng, M * K, L, 40)
    b, eb = rand_mpf_ints(rng, K * N, L, 40)
    rng.shuffle(a)
    a[1], ea[1] = 0, EZ
    ea[2] = EZ if a[2] == 0 else ea[2]
    Ai = [a[i * K:(i + 1) * K] for i in range(M)]
    Ae = [ea[i * K:(i + 1) * K] for i in range(M)]
    Bi = [b[k * N:(k + 1) * N] for k in range(K)]
    Be = [eb[k * N:(k + 1) * N] for k in range(K)]
    for r in range(M):
        Ae[r] = [EZ if v == 0 else e for v, e in zip(Ai[r], Ae[r])]
    for r in range(K):
        Be[r] = [EZ if v == 0 else e for v, e in zip(Bi[r], Be[r])]
    ref = _py_mpf_gemm(Ai, Ae, Bi, Be, L)
    Am = to_planes([v for row in Ai for v in row], L).reshape(L, 1, M, K)
    Aev = np.array(Ae, np.int64).reshape(1, M, K)
    Bm = to_planes([v for row in Bi for v in row], L).reshape(L, 1, K, N)
    Bev = np.array(Be, np.int64).reshape(1, K, N)
    if transA:
        Am, Aev = Am.transpose(0, 1, 3, 2), Aev.transpose(0, 2, 1)
    if transB:
        Bm, Bev = Bm.transpose(0, 1, 3, 2), Bev.transpose(0, 2, 1)
    om, oe = O.mpf_gemm(Am, Aev, Bm, Bev, transA, transB)
    got = from_planes(om.reshape(L, -1))
    for i in range(M):
        for j in range(N):
            assert (got[i * N + j], int(oe[0, i, j])) == ref[i][j]


def test_mpf_gemm_error_in_bits():
    """mathematical check: |C - exact| <= 2^(2-NB) sum_k |a_ik||b_kj| (alignment rounding of both
    operands plus one final rounding), against an mpmath exact product."""
    L, M, K, N = 4, 6, 20, 5
    NB = 32 * L - 2
    rng = random.Random(3)
    a, ea = rand_mpf_ints(rng, M * K, L, 60)
    b, eb = rand_mpf_ints(rng, K * N, L, 60)
    ea = [EZ if v == 0 else e for v, e in zip(a, ea)]
    eb = [EZ if v == 0 else e for v, e in zip(b, eb)]
    om, oe = O.mpf_gemm(to_planes(a, L).reshape(L, 1, M, K), np.array(ea).reshape(1, M, K),
                        to_planes(b, L).reshape(L, 1, K, N), np.array(eb).reshape(1, K, N))
    got = from_planes(om.reshape(L, -1))
    mp.mp.prec = 2000
    val = lambda v, e: mp.mpf(0) if v == 0 else mp.mpf(v) * mp.mpf(2) ** e
    for i in range(M):
        for j in range(N):
            terms = [val(a[i * K + k], ea[i * K + k]) * val(b[k * N + j], eb[k * N + j]) for k in range(K)]
            exact = mp.fsum(terms)
            bound = mp.mpf(2) ** (2 - NB) * sum(abs(x) for x in terms) * 4
            rowmax = max(abs(val(a[i * K + k], ea[i * K + k])) for k in range(K))
            colmax = max(abs(val(b[k * N + j], eb[k * N + j])) for k in range(K))
            bound = max(bound, mp.mpf(2) ** (2 - NB) * K * rowmax * colmax)
            c = val(got[i * N + j], int(oe[0, i, j]))
            assert abs(c - exact) <= bound


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_bgemm_nt_oracle_against_python(mode):
    """batched._bcrt_round rules (sdpvibe/batched.py:200-231) on Python integers"""
    L, nb, M, N, K = 3, 2, 4, 5, 6
    rng = random.Random(9 + mode)
    lim = 1 << (32 * L - 40)
    A = [[[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(M)] for _ in range(nb)]
    Bt = [[[rng.randrange(-lim, lim) for _ in range(K)] for _ in range(N)] for _ in range(nb)]
    Bt[1][2] = [0] * K                                   # an all-zero column
    e_in = np.array([[rng.randrange(-50, 50) for _ in range(N)] for _ in range(nb)], np.int64)
    oexp0 = e_in + np.array([[rng.randrange(-10, 140) for _ in range(N)] for _ in range(nb)])
    Ap = to_planes([v for b in A for r in b for v in r], L).reshape(L, nb, M, K)
    Bp = to_planes([v for b in Bt for r in b for v in r], L).reshape(L, nb, N, K)
    out, oe = O.bgemm_nt(Ap, Bp, e_in, mode, oexp0 if mode == 0 else None)
    got = from_planes(out.reshape(L, -1))
    H = 32 * L - 32
    for b in range(nb):
        exact = [[sum(A[b][i][k] * Bt[b][j][k] for k in range(K)) for j in range(N)] for i in range(M)]
        if mode == 0:
            ex = [int(oexp0[b, j]) for j in range(N)]
        else:
            ex = [int(e_in[b, j]) + max(bitlen_tc(exact[i][j]) for i in range(M)) - H for j in range(N)]
        vals = [[wrap(rnd_half_up(exact[i][j], ex[j] - int(e_in[b, j])), L) for j in range(N)] for i in range(M)]
        if mode == 2:
            emax = max(ex)
            vals = [[wrap(rnd_half_up(vals[i][j], emax - ex[j]), L) for j in range(N)] for i in range(M)]
            ex = [emax] * N
        assert [int(v) for v in oe[b]] == ex
        for i in range(M):
            for j in range(N):
                assert got[(b * M + i) * N + j] == vals[i][j], (b, i, j)


@pytest.mark.parametrize("L,logc", [(4, 10), (8, 60), (24, 80)])
def test_tile_chol_oracle_error_bits(L, logc):
    """chol_tiles/trinv_tiles restatement: L L^T = A and X L = I to the wide precision, and
    |L - L_exact| against an mpmath Cholesky at twice the precision (errors in bits)."""
    from tests.intutil import int_to_words, spd_tile, words_to_int
    BT, LF = 32, L
    WL, F = L + 24, 32 * L
    Ai = spd_tile(BT, logc, L, F)
    W = np.zeros((1, BT, BT, WL), np.uint32)
    for i in range(BT):
        for j in range(BT):
            W[0, i, j] = int_to_words(Ai[i][j], WL)
    Lw, RINV, X = O.tile_chol(W, LF)
    Li = [[words_to_int(Lw[0, i, j]) for j in range(BT)] for i in range(BT)]
    Xi = [[words_to_int(X[0, i, j]) for j in range(BT)] for i in range(BT)]
    assert all(Li[i][j] == 0 for i in range(BT) for j in range(i + 1, BT))
    assert all(Li[i][i] > 0 for i in range(BT))
    one = 1 << F
    # backward error of the factor: max |L L^T - A| in units of 2^-F
    bw = max(abs(sum(Li[i][k] * Li[j][k] for k in range(j + 1)) - (Ai[i][j] << F)) for i in range(BT)
             for j in range(i + 1))
    assert bw.bit_length() <= F + 8, bw.bit_length() - F          # |LL^T - A| <= 2^(8-F)
    # X L = I: error relative to ||X|| ~ cond
    xl = max(abs(sum(Xi[i][k] * Li[k][j] for k in range(BT)) - (one << F if i == j else 0))
             for i in range(BT) for j in range(BT))
    xmax = max(abs(v) for r in Xi for v in r)
    assert xl.bit_length() <= xmax.bit_length() + F + 8
    # forward error against mpmath at 2x precision
    import mpmath as mpm
    mpm.mp.prec = 2 * F + 200
    Am = mpm.matrix([[mpm.mpf(v) / mpm.mpf(2) ** F for v in r] for r in Ai])
    Lr = mpm.cholesky(Am)
    fe = max(abs(mpm.mpf(Li[i][j]) / mpm.mpf(2) ** F - Lr[i, j]) for i in range(BT) for j in range(i + 1))
    cond_bits = int(logc * 3.33)
    assert fe == 0 or float(mpm.log(fe, 2)) <= -F + cond_bits / 2 + 12
