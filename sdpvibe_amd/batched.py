"""Batched multiprecision linear algebra: nb independent blocks of one shape per launch (the
reference's sdpvibe/batched.py interface).

A BMP is a stack of nb fixed-point matrices: limbs (L, nb, R, C) and an exponent that is either one
Python int for the whole stack or one per (block, column), an (nb, C) int64 tensor.

Recursive Cholesky and triangular solves walk a split tree of the dimension (halved at multiples of
the tile size bt, leaves are bt x bt diagonal tiles), so the split points, and with them every
rounding, are those of the reference's recursion; the factor is written in place into one output
array instead of being concatenated level by level.
"""
import torch

from ._lib import call, lib, ptr, ptr_strided, stream
from .limbops import _errbuf, add_same, shift_round
from .mpmat import ctx, headroom_bits, residue_gemm
from .rns import dev


class BMP:
    """stack of nb fixed-point matrices (see the module docstring)"""
    __slots__ = ("limbs", "exp")

    def __init__(self, limbs, exp):
        self.limbs, self.exp = limbs, exp

    @property
    def L(self):
        return self.limbs.shape[0]

    @property
    def shape(self):
        return tuple(self.limbs.shape[1:])

    def expt(self):
        """(nb, C) exponents, materialised when uniform"""
        if isinstance(self.exp, int):
            nb, _, C = self.shape
            return torch.full((nb, C), self.exp, dtype=torch.int64, device=self.limbs.device)
        return self.exp

    def __getitem__(self, idx):
        rows, cols = idx
        e = self.exp if isinstance(self.exp, int) else self.exp[:, cols].contiguous()
        return BMP(self.limbs[:, :, rows, cols].contiguous(), e)

    def t(self):
        """per-block transpose (uniform exponent only)"""
        if not isinstance(self.exp, int):
            raise ValueError("transpose needs a uniform exponent")
        return BMP(self.limbs.transpose(2, 3).contiguous(), self.exp)

    def flat(self):
        return self.limbs.reshape(self.L, -1)


def brescale(A, new_exp, Lout=None, share_exp=False):
    """A re-rounded to new_exp (int, or (nb, C) per block and column), into Lout limbs; the result
    holds a copy of a tensor new_exp, or new_exp itself with share_exp"""
    nb, R, C = A.shape
    Lo = Lout or A.L
    if isinstance(new_exp, int) and isinstance(A.exp, int):
        shift = torch.full((nb * R * C,), new_exp - A.exp, dtype=torch.int64, device=A.limbs.device)
        return BMP(shift_round(A.flat(), shift, Lo).view(Lo, nb, R, C), new_exp)
    if not torch.is_tensor(new_exp):
        new_exp = torch.full((nb, C), int(new_exp), dtype=torch.int64, device=A.limbs.device)
    shift = (new_exp - A.expt()).unsqueeze(1).expand(nb, R, C).reshape(-1)
    return BMP(shift_round(A.flat(), shift, Lo).view(Lo, nb, R, C), new_exp if share_exp else new_exp.clone())


def badd(A, B, alpha=1, out_exp=None):
    """A + alpha B: an exact limb addition when both operands sit at the int exponent out_exp; with
    out_exp 'align', both rounded to one more than the larger exponent of each (block, column) and
    added (one fused kernel)"""
    nb, R, C = A.shape
    L = A.L
    if isinstance(out_exp, int):
        if A.exp != out_exp or B.exp != out_exp:
            raise ValueError("exact addition needs both operands at the output exponent")
        return BMP(add_same(A.flat(), B.flat(), alpha).view(L, nb, R, C), out_exp)
    if out_exp != "align":
        raise ValueError("out_exp must be an int or 'align'")
    out = torch.empty((L, nb, R, C), dtype=torch.int32, device=A.limbs.device)
    e = torch.empty((nb, C), dtype=torch.int64, device=A.limbs.device)
    Al, Bl = A.flat().contiguous(), B.flat().contiguous()
    Ae, Be = A.expt().contiguous(), B.expt().contiguous()
    call("sdpv_badd_align", ptr(Al), ptr(Bl), ptr(Ae), ptr(Be), ptr(out), ptr(e), L, nb, R, C, int(alpha), stream())
    return BMP(out, e)


def _residue_desc(R, X, nb, rows, K):
    """residues of a fixed-point (L, nb, rows, K) limb tensor, all primes; panels of rows j0..j0+jw.
    X may be any view with non-negative strides (a transpose, a row or column slice): the split
    reads it in place through its strides."""
    L = X.shape[0]
    if X.is_contiguous():
        base, plane, ldb, ldr, ldk = X, nb * rows * K, rows * K, K, 1
    else:
        plane, ldb, ldr, ldk = X.stride()
        if min(plane, ldb, ldr, ldk) < 0:
            X = X.contiguous()
            base, plane, ldb, ldr, ldk = X, nb * rows * K, rows * K, K, 1
        else:                                # the storage from X's first entry to its last, flat
            extent = (L - 1) * plane + (nb - 1) * ldb + (rows - 1) * ldr + (K - 1) * ldk + 1
            base = torch.as_strided(X, (extent,), (1,))

    def f(k0, kc, kp, j0, jw):
        r0, rn = (0, rows) if j0 is None else (j0, jw)
        Rh = torch.empty((R.n, kp // 16, nb * rn, 16), dtype=torch.int8, device=X.device)
        Rl = torch.empty_like(Rh)
        # a panel is read in place: rows r0.. of every block, offset r0*ldr into each limb plane
        R.split(base, None, None, L, plane, nb, rn, ldb, ldr, ldk, k0, kc, kp, Rh, Rl, off=r0 * ldr)
        return Rh, Rl
    return f


def _rows(B, r0, r1=None):
    """rows r0..r1 of every block as a view (the exponents are per column, shared)"""
    return BMP(B.limbs[:, :, r0:r1, :], B.exp)


def bgemm_nt(A, Bt, out_exp="col", kc=4096, bt_rowexp=None, bt_cache=None):
    """C[b] = A[b] @ Bt[b]^T for all b (exact, then rounded). A: uniform exponent per block (int or
    (nb, C) tensor read at column 0). Bt: int exponent or per-row exponents bt_rowexp (nb, N).
    out_exp: 'col' (dynamic per (block, column)), 'mat' (dynamic per block), or int (fixed).
    bt_cache: a dict that keeps Bt's residues between calls with the same Bt (its limbs must not
    change while the dict is in use); the residues depend on the limbs only (batched.py:165-197)."""
    nb, M, K = A.shape
    nb2, N, K2 = Bt.shape
    assert nb == nb2 and K == K2
    L = A.L
    R = ctx(L)
    eA = A.exp if isinstance(A.exp, int) else A.exp[:, 0]
    cache = {}
    if not _narrow(M, N, K, L):
        Al, Bl = A.limbs, Bt.limbs                  # views are read through their strides
        same = Al.data_ptr() == Bl.data_ptr() and Al.stride() == Bl.stride() and Al.shape == Bl.shape
        fa = _residue_desc(R, Al, nb, M, K)
        fb = _residue_desc(R, Bl, nb, N, K)

    def fa0(k0, kc_, kp, j0, jw):
        cache["a"] = fa(k0, kc_, kp, None, None)
        return cache["a"]

    def fb0(k0, kc_, kp, j0, jw):
        if same and (j0, jw) == (0, N):
            return cache["a"]
        if bt_cache is None:
            return fb(k0, kc_, kp, j0, jw)
        key = (L, k0, kc_, kp, j0, jw)
        if key not in bt_cache:
            bt_cache[key] = fb(k0, kc_, kp, j0, jw)
        return bt_cache[key]

    eB = Bt.exp if bt_rowexp is None else bt_rowexp
    out = torch.empty((L, nb, M, N), dtype=torch.int32, device=dev)
    if isinstance(out_exp, int) and not _narrow(M, N, K, L):
        # a fixed output exponent: the reconstruction rounds by per-(block, column) shifts alone
        s = _fixed_shift(eA, eB, out_exp, nb, N, bt_cache)
        for j0, jw, acc in residue_gemm(R, fa0, fb0, nb, M, N, K):
            R.crt(acc, nb * M, jw, 0, jw, 2, outm=out, L=L, plane=nb * M * N, ldo=N, jo=j0, Mb=M, scol=s,
                  flag=_errbuf(), scaled=True)
        return BMP(out, out_exp)
    if isinstance(eA, int) and isinstance(eB, int):
        e_in = eA + eB
    else:
        eA_t = eA if torch.is_tensor(eA) else torch.full((nb,), eA, dtype=torch.int64, device=dev)
        eB_t = eB if torch.is_tensor(eB) else torch.full((nb, N), eB, dtype=torch.int64, device=dev)
        if eB_t.dim() == 1:
            eB_t = eB_t[:, None].expand(nb, N)
        e_in = (eA_t[:, None] + eB_t).contiguous()
    exps = None if isinstance(out_exp, int) else torch.empty((nb, N), dtype=torch.int64, device=dev)
    if _narrow(M, N, K, L):
        W, bl = _exact_narrow(A.limbs, Bt.limbs, nb, M, N, K, L)
        _round_exact(W, bl, nb, M, N, e_in, out_exp, L, out, exps)
    else:
        for j0, jw, acc in residue_gemm(R, fa0, fb0, nb, M, N, K):
            _bcrt_round_panel(R, acc, nb, M, N, j0, jw, e_in, out_exp, L, out, exps)
    if isinstance(out_exp, int):
        return BMP(out, out_exp)
    C = BMP(out, exps)
    if out_exp == "mat":
        C = brescale(C, exps.amax(1, keepdim=True).expand(nb, N).contiguous())
    return C


# products with at most NARROW right-hand-side columns and at most NARROW_MK rows times inner length
# are formed directly (sdpv_exact_dot); larger ones go through residues, which are faster there
# (one MI355X, L = 24, one column, M x K: direct 0.084 / 0.104 / 0.106 / 0.164 ms against residues
# 0.135 / 0.134 / 0.132 / 0.135 ms at 32 x 32, 188 x 188, 32 x 1504, 375 x 375)
NARROW = 2
NARROW_MK = 65536


def _narrow(M, N, K, L):
    return N <= NARROW and M * K <= NARROW_MK and L <= 63


def _exact_narrow(Al, Bl, nb, M, N, K, L):
    """the exact products of (L, nb, M, K) and (L, nb, N, K) limb tensors (views with unit stride
    along k are read in place) as 2L + 2 words per entry (entry (b*M + i)*N + j) and bit lengths:
    what the reconstruction's word mode gives for the same product"""
    Al = Al if Al.stride(-1) == 1 else Al.contiguous()
    Bl = Bl if Bl.stride(-1) == 1 else Bl.contiguous()
    NW, E = 2 * L + 2, nb * M * N
    W = torch.empty((NW, E), dtype=torch.int32, device=Al.device)
    bl = torch.empty((E,), dtype=torch.int64, device=Al.device)
    nwork = int(lib().sdpv_exact_dot_work(L, nb, M, N, K))
    work = torch.empty((max(nwork, 1),), dtype=torch.int64, device=Al.device)
    sa, sb = Al.stride(), Bl.stride()
    call("sdpv_exact_dot", ptr_strided(Al), sa[0], sa[1], sa[2], sa[3], ptr_strided(Bl), sb[0], sb[1], sb[2], sb[3],
         L, nb, M, N, K, ptr(W), NW, ptr(bl), ptr(work), nwork, stream())
    return W, bl


def _round_exact(W, bl, nb, M, N, e_in, out_exp, L, out, exps):
    """the roundings of _bcrt_round_panel on exact words: a fixed output exponent, or per
    (block, column) headroom_bits(L) below the column's largest bit length"""
    if isinstance(out_exp, int):
        if isinstance(e_in, int):
            s = torch.full((nb * M * N,), out_exp - e_in, dtype=torch.int64, device=dev)
        else:
            s = (out_exp - e_in)[:, None, :].expand(nb, M, N).reshape(-1).contiguous()
    else:
        ein = e_in if torch.is_tensor(e_in) else torch.full((nb, N), e_in, dtype=torch.int64, device=dev)
        sj = bl.view(nb, M, N).amax(1) - headroom_bits(L)
        exps[:, :] = ein + sj
        s = sj[:, None, :].expand(nb, M, N).reshape(-1).contiguous()
    shift_round(W, s, L, out=out.view(L, nb * M * N))


def _fixed_shift(eA, eB, out_exp, nb, N, cache=None):
    """out_exp - (eA + eB) per (block, column), (nb * N,) int64: the rounding shifts of a product
    with a fixed output exponent. eA: int or (nb,); eB: int, (nb,) or (nb, N). Kept in `cache`
    when eA is an int (the cache of a Bt whose exponents do not change)"""
    key = ("shift", eA, out_exp) if cache is not None and isinstance(eA, int) else None
    if key is not None and key in cache:
        return cache[key]
    if isinstance(eA, int) and isinstance(eB, int):
        s = torch.full((nb * N,), out_exp - (eA + eB), dtype=torch.int64, device=dev)
    else:
        eA_t = eA if torch.is_tensor(eA) else torch.full((nb,), eA, dtype=torch.int64, device=dev)
        eB_t = eB if torch.is_tensor(eB) else torch.full((nb, N), eB, dtype=torch.int64, device=dev)
        if eB_t.dim() == 1:
            eB_t = eB_t[:, None].expand(nb, N)
        s = (out_exp - (eA_t[:, None] + eB_t)).reshape(-1).contiguous()
    if key is not None:
        cache[key] = s
    return s


def _bcrt_round_panel(R, acc, nb, M, N, j0, jw, e_in, out_exp, L, out, exps):
    """_bcrt_round (batched.py:200-231) of one column panel with dynamic output exponents: exact
    products -> L limbs (a fixed output exponent is rounded in bgemm_nt's reconstruction)"""
    rows = nb * M
    Wd, bl, Bq = R.crt_block(acc, jw, 0, jw, rows, True)               # entries (b*M + i)*jw + jj
    ein = e_in if torch.is_tensor(e_in) else torch.full((nb, N), e_in, dtype=torch.int64, device=dev)
    sj = bl.view(nb, M, jw).amax(1) - headroom_bits(L)             # (nb, jw)
    ej = ein[:, j0:j0 + jw] + sj
    s = sj[:, None, :].expand(nb, M, jw).reshape(-1)
    exps[:, j0:j0 + jw] = ej
    if (j0, jw) == (0, N):                                          # the whole width: rounded in place
        shift_round(Wd, s, L, out=out.view(L, nb * M * N))
    else:
        out[:, :, :, j0:j0 + jw] = shift_round(Wd, s, L).view(L, nb, M, jw)


# ------------------------------------------------------------------ batched Cholesky / TRSM
class BTile:
    """a factored diagonal tile: Lmp, the factor (fixed point), and X, its explicit inverse (one
    exponent per block); res: residues of X and Lmp as right operands, kept while the factorisation
    that made the tile reuses them (None otherwise)"""
    __slots__ = ("X", "Lmp", "res")


def bchol_tile(A, LI=24, threads=256):
    """factor and invert the bt x bt tiles A (nb, bt, bt) (fixed point at -(32L - 40)) on the GPU:
    (L as a BMP, BTile) with X rounded to headroom_bits(L) below its largest entry, per block"""
    from .basecase import tile_factor
    from .limbops import bitlen_words
    nb, b, _ = A.shape
    L = A.L
    if A.exp != -(32 * L - 40):
        raise ValueError("tiles enter at the fixed-point exponent")
    Fb = 32 * L
    Lw, Xw, _ = tile_factor(A.limbs, b)
    top = bitlen_words(Xw.reshape(L + LI, -1)).view(nb, b * b).amax(1) - Fb
    tile = BTile()
    tile.res = None
    tile.Lmp = BMP(Lw, A.exp)
    wide = BMP(Xw, torch.full((nb, b), -Fb, dtype=torch.int64, device=Xw.device))
    tile.X = brescale(wide, (top - headroom_bits(L)).unsqueeze(1).expand(nb, b).contiguous(), L)
    return tile.Lmp, tile


class _Node:
    """a diagonal range [lo, lo + n) of the split tree; `head`/`tail` its two halves (None for a
    leaf tile); `first` the index of its first leaf tile, `count` its number of tiles"""
    __slots__ = ("lo", "n", "head", "tail", "first", "count")

    def __init__(self, lo, n, bt, first):
        self.lo, self.n, self.first = lo, n, first
        if n <= bt:
            self.head = self.tail = None
            self.count = 1
            return
        k = (n // 2 + bt - 1) // bt * bt                      # leading part: half, up to a tile multiple
        self.head = _Node(lo, k, bt, first)
        self.tail = _Node(lo + k, n - k, bt, first + self.head.count)
        self.count = self.head.count + self.tail.count

    @property
    def k(self):
        return self.head.n


def _tree(n, bt):
    return _Node(0, n, bt, 0)


# a factorisation keeps the residues of its tiles' X and factor, which the panel solves below them
# use at every level of the split tree (331 leaf solves over 94 tiles at n = 3008), when they take
# at most this many bytes
TILE_RES_BUDGET = 1 << 29


def bchol(A, bt=32, refine=1):
    """Cholesky factors of nb SPD equilibrated fixed-point blocks (nb, n, n), n a multiple of bt:
    (L, list of factored diagonal tiles in tree order). Each internal node factors its head, solves
    for the panel below it (btrsm_right), subtracts the panel's Gram matrix exactly and factors
    the tail; the factor is written into one array as it is produced."""
    nb, n, _ = A.shape
    e = A.exp
    out = torch.zeros((A.L, nb, n, n), dtype=torch.int32, device=A.limbs.device)
    tiles = []
    R = ctx(A.L)
    keep_res = nb * (n // bt) * 2 * 2 * R.n * R.kp(bt) * bt <= TILE_RES_BUDGET

    def factor(node, M):
        lo = node.lo
        if node.head is None:
            if node.n != bt:
                raise ValueError("the dimension must be a multiple of the tile size")
            Lt, t = bchol_tile(M)
            if keep_res:
                t.res = {"X": {}, "L": {}}
            tiles.append(t)
            out[:, :, lo:lo + bt, lo:lo + bt] = Lt.limbs
            return
        k, hi = node.k, lo + node.n
        head, tail = slice(None, k), slice(k, None)
        factor(node.head, _view(M, head, head))
        L11 = _view(BMP(out, e), slice(lo, lo + k), slice(lo, lo + k))
        panel = _right_solve(node.head, _view(M, tail, head), L11, tiles, refine, lo)
        out[:, :, lo + k:hi, lo:lo + k] = panel.limbs
        factor(node.tail, badd(_view(M, tail, tail), bgemm_nt(panel, panel, out_exp=e), -1, out_exp=e))

    factor(_tree(n, bt), A)
    for t in tiles:
        t.res = None
    return BMP(out, e), tiles


def _right_solve(node, B, Lf, tiles, refine, base):
    """X with X Lf^T = B for the node's diagonal block Lf (node offsets relative to `base`)"""
    e = Lf.exp
    if node.head is None:
        t = tiles[node.first]
        cx, cl = (None, None) if t.res is None else (t.res["X"], t.res["L"])
        X = bgemm_nt(B, t.X, out_exp=e, bt_rowexp=t.X.exp, bt_cache=cx)
        for _ in range(refine):                              # one residual correction per step
            resid = badd(B, bgemm_nt(X, t.Lmp, out_exp=e, bt_cache=cl), -1, out_exp=e)
            X = badd(X, bgemm_nt(resid, t.X, out_exp=e, bt_rowexp=t.X.exp, bt_cache=cx), 1, out_exp=e)
        return X
    k = node.k
    head, tail, every = slice(None, k), slice(k, None), slice(None)
    X1 = _right_solve(node.head, _view(B, every, head), _view(Lf, head, head), tiles, refine, base)
    rest = badd(_view(B, every, tail), bgemm_nt(X1, _view(Lf, tail, head), out_exp=e), -1, out_exp=e)
    X2 = _right_solve(node.tail, rest, _view(Lf, tail, tail), tiles, refine, base)
    return BMP(torch.cat([X1.limbs, X2.limbs], 3), e)


def btrsm_right(B, Lf, tiles, bt=32, refine=1):
    """X with X L^T = B, batched, fixed point (B (nb, m, n), L from bchol)"""
    return _right_solve(_tree(B.shape[2], bt), B, Lf, tiles, refine, 0)


def _as_rows(B):
    """B^T per block as the row-exponent operand of bgemm_nt (its column exponents go separately): a
    transposed view, which the residue split reads through its strides"""
    return BMP(B.limbs.transpose(2, 3), 0)


def _apply_tile(X, Lm, B, refine):
    """Lm^-1 B through the explicit tile inverse X, then `refine` residual corrections, with
    per-(block, column) exponents"""
    P = bgemm_nt(X, _as_rows(B), out_exp="col", bt_rowexp=B.expt())
    for _ in range(refine):
        resid = badd(B, bgemm_nt(Lm, _as_rows(P), out_exp="col", bt_rowexp=P.expt()), -1, out_exp="align")
        P = badd(P, bgemm_nt(X, _as_rows(resid), out_exp="col", bt_rowexp=resid.expt()), 1, out_exp="align")
    return P


def _stack_common(P1, P2):
    """[P1; P2] with one exponent per (block, column): both parts re-rounded to the larger"""
    ee = torch.maximum(P1.expt(), P2.expt())
    return BMP(torch.cat([brescale(P1, ee, share_exp=True).limbs, brescale(P2, ee, share_exp=True).limbs], 2), ee)


def _view(A, rows, cols, transpose=False):
    """a sub-block of a uniform-exponent BMP as a view (no copy): bgemm_nt reads views through their
    strides, badd and the tile factorisation copy what they need contiguous"""
    limbs = A.limbs[:, :, rows, cols]
    return BMP(limbs.transpose(2, 3) if transpose else limbs, A.exp)


def _left_solve(node, Lf, B, tiles, refine, upper):
    """L^-1 B (upper=False, forward: head first) or L^-T B (upper=True, backward: tail first)"""
    if node.head is None:
        t = tiles[node.first]
        if upper:
            XT = BMP(t.X.limbs.transpose(2, 3).contiguous(), t.X.exp)
            return _apply_tile(XT, t.Lmp.t(), B, refine)
        return _apply_tile(t.X, t.Lmp, B, refine)
    k = node.k
    head, tail = slice(None, k), slice(k, None)
    if not upper:
        P1 = _left_solve(node.head, _view(Lf, head, head), _rows(B, 0, k), tiles, refine, upper)
        coupled = bgemm_nt(_view(Lf, tail, head), _as_rows(P1), out_exp="col", bt_rowexp=P1.expt())
        P2 = _left_solve(node.tail, _view(Lf, tail, tail), badd(_rows(B, k), coupled, -1, out_exp="align"), tiles,
                         refine, upper)
    else:
        P2 = _left_solve(node.tail, _view(Lf, tail, tail), _rows(B, k), tiles, refine, upper)
        coupled = bgemm_nt(_view(Lf, tail, head, transpose=True), _as_rows(P2), out_exp="col", bt_rowexp=P2.expt())
        P1 = _left_solve(node.head, _view(Lf, head, head), badd(_rows(B, 0, k), coupled, -1, out_exp="align"), tiles,
                         refine, upper)
    return _stack_common(P1, P2)


def btrsm_left(Lf, B, tiles, bt=32, refine=0):
    """P = L^-1 B batched; B (nb, n, m) with per-(block, column) exponents, and so is P"""
    return _left_solve(_tree(B.shape[1], bt), Lf, B, tiles, refine, False)


def btrsm_left_T(Lf, B, tiles, bt=32, refine=1):
    """P = L^-T B batched (L lower, fixed point from bchol)"""
    return _left_solve(_tree(B.shape[1], bt), Lf, B, tiles, refine, True)
