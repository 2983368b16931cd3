"""Exact multiprecision GEMM on CDNA4 INT8 matrix cores via a residue number system (mirror of
sdpvibe/rns.py).

Number format (see mpmat.py): two's-complement integers of L 32-bit limbs in limb planes.
An exact product C = A @ Bt^T of L-limb integers is computed as:
   1. split: every entry is reduced modulo n primes m <= 16257 to a centred residue
      r = 128*h + l (h, l int8), as a byte dot product with a table of 256^j mod m on INT8 MFMA;
   2. products: per prime, three INT8 GEMMs (HH = h.h, SS = (h+l).(h+l), LL = l.l; h + l fits
      int8 because |r| <= 8128), combined as 16256*HH + 128*SS - 127*LL mod m;
   3. CRT: y_g = r_g (M/m_g)^-1 mod m_g (centred; the products can store it directly, their
      coefficients scaled by (M/m_g)^-1), k = round(sum y_g/m_g), Z = [Y, -k] . [digits(M/m_g),
      digits(M)] on INT8 MFMA, carry to words.
Prime count n: the smallest with M > 4 * kmax * 2^(64L-2), so every product with inner dimension
<= kmax is reconstructed exactly and |C| < M/4.

The prime set differs from the reference's (primes <= 32639, four products per prime) so that each
prime costs three INT8 products; the reconstructed integers, and hence every rounded result, are
identical.
"""
import ctypes
import math
import os

import numpy as np
import torch

from ._lib import call, ptr, stream

dev = "cuda"
PMAX = 16257       # largest modulus with centred residues |r| <= 8128: h + l stays in int8
PMIN = 1 << 13
KMAX_LAUNCH = 65536  # inner dimension per product launch (int32 accumulators stay exact)


def _primes_desc(hi, lo):
    """primes in (lo, hi], descending"""
    s = np.ones(hi + 1, bool)
    s[:2] = False
    for p in range(2, int(hi ** 0.5) + 1):
        if s[p]:
            s[p * p::p] = False
    return [int(p) for p in range(hi, lo, -1) if s[p]]


def _sdigits(x, base, n):
    """signed digits in [-base/2, base/2) of a Python int (rns.py:29-35)"""
    h = base // 2
    out = []
    for _ in range(n):
        r = ((x + h) % base) - h
        out.append(r)
        x = (x - r) // base
    assert x == 0
    return out


def gemm_config(M, N, K):
    """(BM, BN) output tile of the product kernel for an M x N x K product (fixed heuristic of the
    library; rns.py:38-54 chose Triton tiles)."""
    mn = min(M, N)
    if mn >= 128:
        return 128, 128
    if mn >= 64:
        return 64, 64
    return 32, 32


# ------------------------------------------------------------------ product tile autotuner
GEMM_CONFIGS = tuple(range(14))   # see sdpv_rns_gemm: tile shapes, stages, three or four products, epilogue
_tuned = {}
TUNE = os.environ.get("SDPV_TUNE", "1") != "0"
# SDPV_GEMM_CONFIG=<c> pins every tuned product to configuration c (measurements of one tile shape)
FORCED = int(os.environ["SDPV_GEMM_CONFIG"]) if os.environ.get("SDPV_GEMM_CONFIG") else None
if FORCED is not None and FORCED not in GEMM_CONFIGS:
    raise ValueError(f"SDPV_GEMM_CONFIG={FORCED}: not one of {GEMM_CONFIGS}")


def _bucket(x):
    """shape class: next power of two (tile choice depends on the size class only)"""
    return 1 << max(0, int(x - 1).bit_length())


# square output tiles: the syrk mode's lower-tile test needs BM = BN. Timed in the syrk mode at
# N = 3000 (GramAcc), none beats the 128 x 128 four-product tile of the size rule, so syrk products
# are not tuned (no first-iteration tuning cost); a caller may still pin one of these.
SYRK_CONFIGS = (0, 1, 2, 3, 6)


def tuned_config(R, nb, M, N, kp):
    """fastest product tile configuration for this shape class, measured once per process on
    scratch residues (every configuration computes identical residues); -1 (size rule) when
    tuning is off or the product is small"""
    if not TUNE or nb * M * N * kp < (1 << 24):
        return -1
    if FORCED is not None:
        return FORCED
    key = (R.L, _bucket(nb), _bucket(M), _bucket(N), _bucket(kp))
    if key not in _tuned:
        _tuned[key] = _measure(R, nb, M, N, kp)
    return _tuned[key]


def _measure(R, nb, M, N, kp, reps=3):
    # every prime when the scratch operands stay small (the launch's wave count decides between tiles)
    G = R.n if R.n * nb * (2 * kp * (M + N) + 2 * M * N) <= (1 << 29) else min(R.n, 32)
    dev_ = R.m32.device
    gen = torch.Generator(device=dev_)
    gen.manual_seed(0)
    mk = lambda rows: torch.randint(-64, 64, (G, kp // 16, nb * rows, 16), dtype=torch.int8, device=dev_,
                                    generator=gen)
    Ah, Al, Bh, Bl = mk(M), mk(M), mk(N), mk(N)
    acc = torch.empty((G, nb * M, N), dtype=torch.int16, device=dev_)
    st = torch.cuda.current_stream()
    launch = lambda cfg: call("sdpv_rns_gemm", ptr(Ah), ptr(Al), ptr(Bh), ptr(Bl), G, nb, M, N, kp, ptr(acc),
                              nb * M * N, ptr(R.m32), ptr(R.minv), ptr(R.kcoef), 0, 0, cfg, stream())
    for cfg in GEMM_CONFIGS:                          # warm every configuration (attributes, clocks)
        launch(cfg)
    best_t = {}
    for _ in range(2):                                # two passes, the faster of each: clock drift
        for cfg in GEMM_CONFIGS:
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(st)
            for _ in range(reps):
                launch(cfg)
            ev1.record(st)
            ev1.synchronize()
            t = ev0.elapsed_time(ev1)
            best_t[cfg] = min(t, best_t.get(cfg, t))
    return min(GEMM_CONFIGS, key=lambda c: best_t[c])


def row_major(R):
    """chunk-major residues (G, kp/16, rows, 16) -> (G, rows, kp)"""
    G, nc, rows, _ = R.shape
    return R.permute(0, 2, 1, 3).reshape(G, rows, nc * 16).contiguous()


def chunk_major(R):
    """(G, rows, kp) residues -> the kernels' chunk-major layout (G, kp/16, rows, 16)"""
    G, rows, kp = R.shape
    return R.reshape(G, rows, kp // 16, 16).permute(0, 2, 1, 3).contiguous()


class RNS:
    """Moduli and tables for exact products of L-limb integers with inner dimension <= kmax
    (rns.py:163-191)."""
    KPAD = 64

    def __init__(self, L, kmax=1 << 17, G=16, device=None):
        self.L, self.G, self.NB, self.kmax = L, G, 4 * L, kmax
        device = device if device is not None else torch.device(dev, torch.cuda.current_device())
        bound = kmax * (1 << (32 * L - 1)) ** 2              # |C| < bound
        mods, M = [], 1
        for p in _primes_desc(PMAX, PMIN):
            if M > 4 * bound:
                break
            mods.append(p)
            M *= p
        if M <= 4 * bound:
            raise ValueError(f"L={L}: not enough primes below {PMAX}")
        self.mods, self.M, self.n = mods, M, len(mods)
        n = self.n
        self.npad = -(-(n + 1) // 32) * 32                 # column n of Ct: digits of M
        self.NBp = 32 * (-(-4 * L // 32))
        t = lambda a, dt: torch.tensor(a, dtype=dt, device=device)
        self.m32 = t(mods, torch.int32)
        self.minv = t([1.0 / m for m in mods], torch.float64)
        cen = lambda v, m: v - m if v > m // 2 else v
        T = np.zeros((self.npad, self.NBp), np.int64)
        for i, m in enumerate(mods):
            T[i, :4 * L] = [cen(pow(256, j, m), m) for j in range(4 * L)]
        T_lo = ((T + 64) % 128) - 64
        T_hi = (T - T_lo) // 128
        self.Thi = t(T_hi, torch.int8).contiguous()
        self.Tlo = t(T_lo, torch.int8).contiguous()
        # every byte but the top one enters as u - 128, the top one signed: C1 restores the bias
        self.C1 = t([(128 * int(T[i, :4 * L - 1].sum())) % m for i, m in enumerate(mods)], torch.int32)
        self.kcoef = t([[16256 % m, 128 % m, (-127) % m] for m in mods], torch.int32).contiguous()
        inv = [pow(M // m % m, -1, m) for m in mods]
        # coefficients scaled by (M/m)^-1: the product then stores the CRT's y_g directly
        self.kcoef_crt = t([[16256 * v % m, 128 * v % m, (-127) * v % m] for m, v in zip(mods, inv)],
                           torch.int32).contiguous()
        # CRT in base 256
        self.E = (math.ceil(M.bit_length() / 8) + 2 + 7) // 8 * 8
        self.Epad = -(-self.E // 32) * 32
        self.inv32 = t(inv, torch.int32)
        Ct = np.zeros((self.Epad, self.npad), np.int64)
        for i, m in enumerate(mods):
            Ct[:self.E, i] = _sdigits(M // m, 256, self.E)
        Md = np.zeros(self.Epad, np.int64)
        Md[:self.E] = _sdigits(M, 256, self.E)
        Ct[:, n] = Md                                      # the CRT subtracts k*M with the digit products
        self.Ct = t(Ct, torch.int8).contiguous()
        self.Md = t(Md, torch.int32)

    # ---- step 1: residues of rows x kc entries (limbs) -> (G, rows, kp) hi, lo
    def kp(self, kc):
        return -(-kc // self.KPAD) * self.KPAD

    def residues(self, X, k0, kc, g0=0, G=None):
        """X: (L, rows, K) int32 limb planes. Residues of X[:, :, k0:k0+kc] modulo primes
        g0..g0+G as (G, rows, kp) int8 hi / lo (128*hi + lo = centred residue), inner dimension
        zero-padded to kp = kc rounded up to a multiple of 64 (rns.py:196-214)."""
        L, rows, K = X.shape
        assert L == self.L
        G = self.n - g0 if G is None else G
        kp = self.kp(kc)
        X = X.contiguous()
        Rh = torch.empty((G, kp // 16, rows, 16), dtype=torch.int8, device=X.device)
        Rl = torch.empty_like(Rh)
        self.split(X, None, None, L, rows * K, 1, rows, 0, K, 1, k0, kc, kp, Rh, Rl, g0, G)
        return row_major(Rh), row_major(Rl)

    def split(self, X, e, Erow, Lin, plane, nb, rows, ldb, ldr, ldk, k0, kc, kp, Rh, Rl, g0=0, G=None,
              flag=None, off=0):
        """Fused rounding shift + residue split (see sdpv_rns_split): entry (b, r, k0 + k) at
        off + b*ldb + r*ldr + (k0+k)*ldk of every limb plane of X (Lin limbs) and of e, shifted by
        Erow[b*rows + r] - e (none when e is None) into self.L limbs."""
        G = self.n - g0 if G is None else G
        # host-side extent checks: the kernel trusts these sizes
        nrows = nb * rows
        if not (0 <= g0 and g0 + G <= self.n and kp % self.KPAD == 0 and 0 <= kc <= kp):
            raise ValueError("split: prime range or padded width out of range")
        for Rx in (Rh, Rl):
            if tuple(Rx.shape) != (G, kp // 16, nrows, 16) or not Rx.is_contiguous() or Rx.dtype != torch.int8:
                raise ValueError(f"split: residue planes must be contiguous int8 {(G, kp // 16, nrows, 16)}, "
                                 f"got {tuple(Rx.shape)}")
        if nrows and kc:
            last = off + (nb - 1) * ldb + (rows - 1) * ldr + (k0 + kc - 1) * ldk
            if min(ldb, ldr, ldk, k0, off) < 0 or last >= plane or (Lin - 1) * plane + last >= X.numel():
                raise ValueError("split: entry offsets exceed the limb planes")
            if e is not None and (e.numel() < plane or Erow is None or Erow.numel() < nrows):
                raise ValueError("split: exponent arrays too small")
        Xp, ep = ptr(X), ptr(e)
        if off:                                # the entry point reads every plane from X + off
            Xp = ctypes.c_void_p(Xp.value + off * X.element_size())
            ep = None if e is None else ctypes.c_void_p(ep.value + off * e.element_size())
        call("sdpv_rns_split", Xp, ep, ptr(Erow), Lin, self.L, plane, nb, rows, ldb, ldr, ldk, k0, kc,
             kp, ptr(self.Thi), ptr(self.Tlo), self.NBp, ptr(self.C1), ptr(self.m32),
             ptr(self.minv), g0, G, ptr(Rh), ptr(Rl), ptr(flag), stream())

    # ---- step 2: accumulate residues of A @ Bt^T
    def product(self, Ah, Al, Bh, Bl, acc, nb, M, N, ld_g, g0=0, accumulate=False, syrk=False, config=None,
                scaled=False):
        """residues in the kernels' chunk-major layout (G, kp/16, rows, 16). scaled: store the CRT
        coefficients y_g = centred(residue * (M/m_g)^-1 mod m_g) instead of residues in [0, m)
        (crt(..., scaled=True) then skips that step)."""
        if Ah.dim() != 4:
            raise ValueError("product: residues must be in the chunk-major layout (G, kp/16, rows, 16)")
        G, kp = Ah.shape[0], Ah.shape[1] * 16
        for X, rows in ((Ah, nb * M), (Al, nb * M), (Bh, nb * N), (Bl, nb * N)):
            if tuple(X.shape) != (G, kp // 16, rows, 16) or not X.is_contiguous():
                raise ValueError(f"product: residues must be contiguous {(G, kp // 16, rows, 16)}, "
                                 f"got {tuple(X.shape)}")
        if g0 < 0 or g0 + G > self.n or ld_g < nb * M * N or acc.numel() < (g0 + G - 1) * ld_g + nb * M * N:
            raise ValueError("product: accumulator too small for the primes and output block")
        mode = (1 if accumulate else 0) | (2 if syrk else 0) | (4 if scaled else 0)
        if syrk and M != N:
            raise ValueError("product: the syrk mode needs M == N")
        if config is None:
            config = -1 if syrk else tuned_config(self, nb, M, N, kp)
        elif syrk and config not in (-1,) + SYRK_CONFIGS:
            raise ValueError(f"product: configuration {config} has no square tile for the syrk mode")
        call("sdpv_rns_gemm", ptr(Ah), ptr(Al), ptr(Bh), ptr(Bl), G, nb, M, N, kp, ptr(acc), ld_g,
             ptr(self.m32), ptr(self.minv), ptr(self.kcoef_crt if scaled else self.kcoef), g0, mode, config,
             stream())

    def accumulate(self, Ahi, Alo, Bhi, Blo, acc, g0=0):
        """acc[g0:g0+G] += residues of A @ B^T for the G moduli whose residues are given in the
        (G, rows, kp) layout of residues() (rns.py:217-224)"""
        G, M, KC = Ahi.shape
        N = Bhi.shape[1]
        assert acc.is_contiguous() and acc.shape[1:] == (M, N)
        Ah, Al, Bh, Bl = (chunk_major(x) for x in (Ahi, Alo, Bhi, Blo))
        self.product(Ah, Al, Bh, Bl, acc, 1, M, N, M * N, g0, accumulate=True)

    # ---- step 3: CRT of a block of entries
    def crt(self, acc, rows, N, j0, w, mode, ld_g=None, W=None, bitlen=None, outm=None, oute=None, L=None,
            plane=0, ldo=0, jo=0, Mb=1, erow=None, ecol=None, scol=None, flag=None, scaled=False):
        ld_g = acc.stride(0) if ld_g is None else ld_g
        if acc.numel() < (self.n - 1) * ld_g + (rows - 1) * N + j0 + w or j0 + w > N:
            raise ValueError("crt: block exceeds the residue accumulator")
        Bq = rows * w
        if mode == 0 and (W is None or W.numel() < (self.E // 4) * Bq or bitlen is None or bitlen.numel() < Bq):
            raise ValueError("crt: word outputs too small")
        if mode != 0 and Bq and (outm is None or outm.numel() < ((L or self.L) - 1) * plane + (rows - 1) * ldo
                                 + jo + w):
            raise ValueError("crt: limb output too small")
        call("sdpv_rns_crt", ptr(acc), self.n, ld_g, rows, N, j0, w, ptr(self.m32), ptr(self.minv),
             ptr(self.inv32), ptr(self.Ct), self.E, self.npad, mode | (4 if scaled else 0), ptr(W), ptr(bitlen),
             ptr(outm), ptr(oute), L or self.L, plane, ldo, jo, Mb, ptr(erow), ptr(ecol), ptr(scol),
             ptr(flag), stream())

    def crt_block(self, acc, N, j0, w, rows, scaled=False):
        """acc: (n, rows_total, N) int16 residues in [0, m) (scaled: the product's CRT coefficients);
        block = rows < `rows`, columns j0..j0+w; entries q = i*w + jj. Returns (W (E/4, Bq) int32
        words, bitlen (Bq,), Bq) (rns.py:227-245)."""
        Bq = rows * w
        acc = acc.contiguous()
        W = torch.empty((self.E // 4, Bq), dtype=torch.int32, device=acc.device)
        bl = torch.empty((Bq,), dtype=torch.int64, device=acc.device)
        self.crt(acc, rows, N, j0, w, 0, W=W, bitlen=bl, scaled=scaled)
        return W, bl, Bq
