/*
 * sdpvibe_hip.h -- C ABI of the MI355X (gfx950) multiprecision linear-algebra core.
 *
 * Number format (shared by every entry point):
 *   - a multiword integer of L limbs is stored in limb-plane layout: limb k of entry e lives at
 *     X[k * plane + e] (int32, little-endian limbs, two's complement in the top limb);
 *   - exponents are int64; an MPF value is m * 2^e, zero has m = 0 and e = SDPV_EZ.
 *
 * Conventions:
 *   - every pointer is a device pointer unless stated otherwise; the caller owns every buffer;
 *   - `stream` is a hipStream_t; every kernel is enqueued on it and no entry point synchronises
 *     with the host, so calls can be recorded into a hipGraph;
 *   - return value: SDPV_OK, or a negative SDPV_E* code (bad argument, unsupported size, launch
 *     failure). Numeric overflow of a rounding step is reported through a caller-supplied int32
 *     device flag instead (sticky: kernels only ever set it to 1).
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * AntoineTilloy/SDPviBe sources).
 */
#ifndef SDPVIBE_HIP_H
#define SDPVIBE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDPV_OK 0
#define SDPV_EARG (-1)      /* bad argument (null pointer, negative size, bad mode) */
#define SDPV_ESIZE (-2)     /* size outside what the kernels support (e.g. L too large) */
#define SDPV_ELAUNCH (-3)   /* kernel launch failed */

#define SDPV_EZ (-(1LL << 40)) /* exponent of zero (sdpvibe/mpf.py:21) */

/* Library version; bumped whenever an entry point changes. */
int sdpv_version(void);

/* Hash of this header and of the kernel sources the library was built from (the build passes it
 * in). Loaders compare it with the sources they ship, so a stale binary is refused. */
uint64_t sdpv_source_hash(void);

/* Page-locked host memory of exactly `bytes` bytes (hipHostMalloc), NULL when it cannot be had;
 * release it with sdpv_host_free. The streaming store keeps the SDP's large per-chunk arrays there
 * and copies them to and from the device asynchronously. Replaces the cudaHostRegister of
 * torch-allocated buffers in Store._alloc_host (sdpvibe/store.py:52-65). */
void* sdpv_host_alloc(int64_t bytes);
int sdpv_host_free(void* p);

/* ---------------------------------------------------------------- limb bookkeeping */

/* out = round_half_up(X / 2^s) per entry as Lout limbs (s <= 0: exact left shift).
 * X: (Lin, B) planes, s: (B,), out: (Lout, B). flag: per_entry ? (B,) int32 of 0/1
 * : one int32 set to 1 when any entry overflows Lout limbs.
 * Replaces limbops.shift_round / _shift_round_kernel (sdpvibe/limbops.py:173-185, :32-62). */
int sdpv_shift_round(const int32_t* X, int64_t Lin, int64_t B, const int64_t* s, int32_t* out,
                     int64_t Lout, int32_t* flag, int32_t per_entry, void* stream);

/* out = A + sgn * Bm (mod 2^(32 L)), sgn = +1 or -1. (L, B) planes.
 * Replaces limbops.add_same / _add_kernel (sdpvibe/limbops.py:188-193, :66-81). */
int sdpv_add_same(const int32_t* A, const int32_t* Bm, int32_t* out, int64_t L, int64_t B,
                  int32_t sgn, void* stream);

/* out[e] = bit length of v (of ~v when v < 0), v = X[:, e].
 * Replaces limbops.bitlen_words / _bitlen_kernel (sdpvibe/limbops.py:213-218, :197-210). */
int sdpv_bitlen(const int32_t* X, int64_t L, int64_t B, int64_t* out, void* stream);

/* Batched aligned add: entry (b, r, c) of A and Bm (each (L, nb, rows, cols) planes) with
 * column exponents EA[b*cols + c], EB[...]: ex = max(ea, eb) + 1, both operands rounded
 * (half up) to 2^ex independently, then out = a' + sgn * b' (wraps); EO[b*cols + c] = ex.
 * Replaces batched.badd(..., 'align') / _badd_align_kernel (sdpvibe/batched.py:56-92). */
int sdpv_badd_align(const int32_t* A, const int32_t* Bm, const int64_t* EA, const int64_t* EB,
                    int32_t* out, int64_t* EO, int64_t L, int64_t nb, int64_t rows, int64_t cols,
                    int32_t sgn, void* stream);

/* Exact products of a batch of fixed-point matrices with a few vectors, without residues: for every
 * entry e = (b * M + i) * N + j, S = sum_k A[b, i, k] * B[b, j, k] over L-limb two's-complement
 * operands (limb u of A[b, i, k] at A[u * planeA + b * ldAb + i * ldAr + k * ldAk], likewise B),
 * written exactly as NW >= 2L + 2 words W[w * E + e] (E = nb * M * N), with its bit length (of ~S
 * when S < 0) in bitlen[e]: the words and bit lengths that sdpv_rns_crt's word mode returns for the
 * same product. K beyond 256 is split over waves, whose column sums go through `work`
 * (sdpv_exact_dot_work(...) int64 entries; 0 when K <= 256). No counterpart in the reference, which
 * always goes through residues: the direct path for products with one or two right-hand-side
 * columns (triangular solves with one vector), where the residue path's fixed costs dominate.
 * L <= 63. */
int64_t sdpv_exact_dot_work(int64_t L, int64_t nb, int64_t M, int64_t N, int64_t K);
int sdpv_exact_dot(const int32_t* A, int64_t planeA, int64_t ldAb, int64_t ldAr, int64_t ldAk,
                   const int32_t* B, int64_t planeB, int64_t ldBb, int64_t ldBr, int64_t ldBk,
                   int64_t L, int64_t nb, int64_t M, int64_t N, int64_t K, int32_t* W, int64_t NW,
                   int64_t* bitlen, int64_t* work, int64_t nwork, void* stream);

/* ---------------------------------------------------------------- per-entry floats (MPF) */

/* top[e] = e + bitlen(m) (bitlen of ~m for m < 0), or SDPV_EZ when e == SDPV_EZ.
 * Replaces mpf._top (sdpvibe/mpf.py:348-351). */
int sdpv_mpf_top(const int32_t* m, const int64_t* e, int64_t L, int64_t n, int64_t* top,
                 void* stream);

/* Exact product, magnitude rounded half up to NB = 32L-2 bits; B may be one broadcast entry.
 * Replaces mpf.mul / mpf_mul kernel (sdpvibe/mpf.py:316-323, :58-89). */
int sdpv_mpf_mul(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB,
                 int32_t* O, int64_t* eO, int64_t L, int64_t n, int32_t bscalar, void* stream);

/* O = A + sgn * Bm, normalised. Replaces mpf.add / mpf_add (sdpvibe/mpf.py:326-331, :92-124). */
int sdpv_mpf_add(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB,
                 int32_t* O, int64_t* eO, int64_t L, int64_t n, int32_t sgn, void* stream);

/* Normalise LI-limb two's-complement integers X (LI, n) with exponents eI to L limbs.
 * Replaces mpf.normalize / mpf_norm (sdpvibe/mpf.py:307-313, :127-140). */
int sdpv_mpf_norm(const int32_t* X, const int64_t* eI, int64_t LI, int32_t* O, int64_t* eO,
                  int64_t L, int64_t n, void* stream);

/* ---------------------------------------------------------------- exact RNS GEMM */

/* Operand geometry shared by the split entry points: entry (b, r, k) of a batch of nb matrices
 * with `rows` rows lives at b*ldb + r*ldr + k*ldk inside every limb plane (plane stride
 * `plane`) and inside the exponent array e. */

/* Row alignment exponents: Erow[b*rows + r] = max_k top(b, r, k) - (32 L - 2) over the entries
 * k in [0, K) with e != SDPV_EZ; SDPV_EZ when there is none (the reference then uses 0, which no
 * product can observe: every entry of such a row is zero). Rows contiguous in k: one launch, one
 * wave per row. Other layouts: one launch when there are enough groups of 64 rows, else a fill
 * and a launch split over k that meets by atomic maximum. Replaces the exponent half of
 * mpf._align_rows (sdpvibe/mpf.py:390-398). */
int sdpv_row_exp(const int32_t* X, const int64_t* e, int64_t L, int64_t plane, int64_t nb,
                 int64_t rows, int64_t K, int64_t ldb, int64_t ldr, int64_t ldk, int64_t* Erow,
                 void* stream);

/* Residue split. For every entry (b, r, k0 + k), k < kc: the value is first shift-rounded
 * (as sdpv_shift_round) by s = Erow[b*rows + r] - e[entry] (s = 2^30 when e == SDPV_EZ;
 * s = 0 for every entry when e == NULL) from L to Lout limbs, then reduced modulo the primes
 * g0 .. g0+G-1 to the centred residue r = 128*hi + lo (lo in [-64, 63]).
 * Rh, Rl: (G, kp/16, nb*rows, 16) int8, chunk-major: byte k of row r of prime g at
 * ((g*(kp/16) + k/16)*nb*rows + r)*16 + k%16; columns kc <= k < kp hold residue 0.
 * Tables (built by the caller for Lout): Thi/Tlo (npad, NBp) int8 with NBp = 32*ceil(4 Lout/32),
 * T = 128*Thi + Tlo the centred 256^j mod m; C1 (n,) int32 = 128 * sum_{j < 4 Lout - 1} T_j mod m
 * (the bias of every byte but the top one, which enters signed); mods (n,) int32, minv (n,)
 * double = 1/m.
 * flag: optional int32 set to 1 when the shifted value overflows Lout limbs.
 * Replaces RNS.residues / _residue_kernel (sdpvibe/rns.py:196-214, :67-103) fused with the
 * shift_round of mpf._align_rows / GramAcc.add (sdpvibe/mpf.py:390-398, :574-576). */
int sdpv_rns_split(const int32_t* X, const int64_t* e, const int64_t* Erow, int64_t L,
                   int64_t Lout, int64_t plane, int64_t nb, int64_t rows, int64_t ldb,
                   int64_t ldr, int64_t ldk, int64_t k0, int64_t kc, int64_t kp,
                   const int8_t* Thi, const int8_t* Tlo, int64_t NBp, const int32_t* C1,
                   const int32_t* mods, const double* minv, int64_t g0,
                   int64_t G, int8_t* Rh, int8_t* Rl, int32_t* flag, void* stream);

#define SDPV_GEMM_ACCUMULATE 1 /* acc = (acc + product) mod m instead of acc = product mod m */
#define SDPV_GEMM_SYRK_LOWER 2 /* B is A (M == N); only tiles on or below the diagonal */
#define SDPV_GEMM_CENTRED 4    /* store residues centred in [-(m-1)/2, (m-1)/2] instead of [0, m) */

/* Modular products of residues, per prime: for gi < G, b < nb, with a = 128*Ah + Al and
 * b = 128*Bh + Bl the centred residues,
 *   acc[(g0+gi)*ld_g + (b*M + i)*N + j] = s_g * (sum_k a[gi, b*M+i, k] * b[gi, b*N+j, k]) mod m_g
 * (g = g0+gi) stored as int16 in [0, m), or centred with SDPV_GEMM_CENTRED.
 * Ah/Al: (G, kp/16, nb*M, 16), Bh/Bl: (G, kp/16, nb*N, 16) int8, chunk-major as written by
 * sdpv_rns_split, kp a multiple of 64, kp <= 65536.
 * kcoef: (n, 3) int32 = (16256, 128, -127) * s_g mod m: ra*rb = 16256*HH + 128*SS - 127*LL.
 * s_g = 1 gives the residues of the product; s_g = (M/m_g)^-1 mod m_g with SDPV_GEMM_CENTRED
 * gives the CRT coefficients that sdpv_rns_crt takes with SDPV_CRT_SCALED.
 * Replaces RNS.accumulate / _gemm_kernel (sdpvibe/rns.py:217-224, :107-137),
 * _bgemm_kernel (sdpvibe/batched.py:97-127) and _syrk_lower_kernel (sdpvibe/batched.py:131-162). */
/* config: -1 picks the tile by size; with SDPV_GEMM_SYRK_LOWER only -1 and the square tiles
 * 0, 1, 2, 3 and 6 are accepted; 0..5 select
 * 128x128 (2 stages), 128x128 (3 stages), 64x64, 32x32, 256x128 and 128x256 tiles with three
 * INT8 products per prime (the sum plane h + l formed in registers); 6 and 7 are the 128x128 and
 * 128x256 tiles with four products (HH, LL, HL + LH; no bytewise adds); 8 and 9 the 128x256 tile
 * with 3 stages, four and three products; 10 and 11 the 128x64 and 64x128 tiles with four
 * products (fewer registers, more workgroups per CU), 12 the 64x128 tile in 3 stages of 32 bytes of
 * k (36 KB of LDS, four workgroups per CU); 13 is configuration 12 with its epilogue staged
 * through LDS, so that each lane stores 16 contiguous bytes (full tiles, N a multiple of 8).
 * Every configuration computes the same residues (an autotuner picks by time). */
int sdpv_rns_gemm(const int8_t* Ah, const int8_t* Al, const int8_t* Bh, const int8_t* Bl,
                  int64_t G, int64_t nb, int64_t M, int64_t N, int64_t kp, int16_t* acc,
                  int64_t ld_g, const int32_t* mods, const double* minv, const int32_t* kcoef,
                  int64_t g0, int32_t mode, int32_t config, void* stream);

#define SDPV_CRT_WORDS 0     /* W (E/4, B) int32 two's complement + bitlen (B,) */
#define SDPV_CRT_NORMALIZE 1 /* MPF output: L limbs with NB bits, exponent ein + shift */
#define SDPV_CRT_ROUND 2     /* fixed-exponent output: round_half_up(value / 2^s) to L limbs */
#define SDPV_CRT_SCALED 4    /* flag: acc holds y = residue * (M/m)^-1 mod m, centred */

/* CRT reconstruction of the exact integers held as residues acc[g*ld_g + i*N + j]
 * (g < n, 0 <= value residue < m_g; with SDPV_CRT_SCALED in mode the centred
 * residue * (M/m_g)^-1 mod m_g) for the block of entries (i, j0 + jj), i < rows, jj < w;
 * entry index q = i*w + jj. Tables: mods, minv (1/m double), inv ((M/m)^-1 mod m),
 * Ct (Epad, npad) int8, Epad = 32*ceil(E/32), npad > n: column g < n holds the base-256 signed
 * digits of M/m_g, column n the digits of M, the rest and rows >= E are zero.
 * mode SDPV_CRT_WORDS:     W[t*Bq + q] (t < E/4), bitlen[q]              (Bq = rows * w)
 * mode SDPV_CRT_NORMALIZE: outm[k*plane + o], oute[o], o = i*ldo + jo + jj, with
 *                          ein = erow[i] + ecol[c], c = (i / Mb)*ldo + jo + jj (NULL arrays = 0)
 * mode SDPV_CRT_ROUND:     outm[k*plane + o] = round_half_up(value / 2^scol[c]); overflow -> flag
 * Replaces RNS.crt_block (sdpvibe/rns.py:227-245) = crt_pre + _crt_dot_kernel + crt_carry
 * (sdpvibe/limbops.py:86-152, :221-239) fused with mpf.normalize / shift_round
 * (sdpvibe/mpf.py:440-450, sdpvibe/mpmat.py:205-236, sdpvibe/batched.py:200-231). */
int sdpv_rns_crt(const int16_t* acc, int64_t n, int64_t ld_g, int64_t rows, int64_t N,
                 int64_t j0, int64_t w, const int32_t* mods, const double* minv,
                 const int32_t* inv, const int8_t* Ct, int64_t E, int64_t npad,
                 int32_t mode, int32_t* W, int64_t* bitlen, int32_t* outm, int64_t* oute,
                 int64_t L, int64_t plane, int64_t ldo, int64_t jo, int64_t Mb, const int64_t* erow,
                 const int64_t* ecol, const int64_t* scol, int32_t* flag, void* stream);

/* ---------------------------------------------------------------- tile base case */

/* Batched Cholesky factor and inverse of nb BT x BT tiles (BT <= 64) in wide fixed point.
 * A: (L, nb, BT, BT) limb planes of equilibrated SPD blocks at exponent -(32L - 40).
 * The tiles are widened exactly to WL = L + 24 words with F = 32L fractional bits, factored
 * (pivot d^-1/2 by a double seed and precision-doubling Newton, truncated products) and inverted
 * row by row, as the reference's kernels do.
 * Lout: (L, nb, BT, BT) the factor rounded (half up) back to A's exponent, upper triangle zero.
 * Xw:   (L + 24, nb, BT, BT) the inverse L^-1 in wide fixed point (exponent -32L).
 * work: device scratch of nb * BT * (3 * BT + 1) * (L + 24) uint32 words. On return its first
 *       nb * BT * BT * (L + 24) words hold the wide factor (nb, BT, BT, L + 24), upper triangle zero,
 *       and the words from 3 * nb * BT * BT * (L + 24) on the pivot reciprocals (nb, BT, L + 24): the
 *       reference's Tile.Lw and Tile.RINV.
 * flag: optional int32 set to 1 when the narrowing to L limbs overflows.
 * Replaces chol_tiles + trinv_tiles (sdpvibe/basecase.py:159-226) and the widening/narrowing of
 * batched.bchol_tile (sdpvibe/batched.py:243-257). */
int sdpv_tile_chol(const int32_t* A, int64_t L, int64_t nb, int64_t BT, int32_t* Lout, int32_t* Xw,
                   uint32_t* work, int32_t* flag, void* stream);

/* Forward substitution L y = v with one factored tile, in place, for nvec vectors.
 * Lw: (BT, BT, WL) wide factor, RINV: (BT, WL) pivot reciprocals (both from sdpv_tile_chol's work),
 * V: (nvec, BT, WL) wide fixed-point vectors (F = 32 LF fractional bits). BT <= 32.
 * Replaces trsv_warp (sdpvibe/basecase.py:228-245). */
int sdpv_tile_trsv(const uint32_t* Lw, const uint32_t* RINV, uint32_t* V, int64_t nvec, int64_t BT, int64_t WL,
                   int64_t LF, void* stream);

/* ---------------------------------------------------------------- FP64 step-length pieces */

/* out[i] = MPF.to_float of entry i (m: (L, n) planes, e: (n,)) after multiplying it by
 * 2^-shift[i / per] (shift may be null): the top two limbs as a double times
 * 2^clamp(e - shift + 32(L - 2), -1100, 1100); zeros give 0.
 * Replaces MPF.shift(...).to_float() (sdpvibe/mpf.py:194-197, :210-215). */
int sdpv_mpf_to_double(const int32_t* m, const int64_t* e, int64_t L, int64_t n, const int64_t* shift,
                       int64_t per, double* out, void* stream);

/* Exact conversion of n doubles to normalised MPF (L >= 2): mantissa * 2^(32L - 55), zero -> EZ.
 * Replaces solver._mpf_from_float (sdpvibe/solver.py:1019-1025). */
int sdpv_mpf_from_double(const double* v, int64_t n, int64_t L, int32_t* m, int64_t* e, void* stream);

/* Eigenvalues (ascending, w (nb, n)) and optionally eigenvectors (columns of V (nb, n, n)) of the
 * symmetrised (A + A^T)/2 of nb float64 n x n matrices (n <= SDPV_SYMEIG_NMAX), their Frobenius
 * norms (fro (nb,), may be null) and the Frobenius norm of what is left off the diagonal when the
 * iteration stops (off (nb,), may be null): each returned eigenvalue is within off of an exact
 * eigenvalue of the float64 matrix (Weyl). work: 2 * nb * m * m doubles, m = n rounded up to even.
 * Parallel cyclic Jacobi. Replaces numpy.linalg.eigvalsh / eigh in min_eig and refine_min_eig
 * (sdpvibe/solver.py:1039, :1077). */
#define SDPV_SYMEIG_NMAX 4096
int sdpv_symeig(const double* A, int64_t nb, int64_t n, double* work, double* w, double* V, double* fro,
                double* off, void* stream);

/* sdpv_symeig without the work argument: the work area lives in a device array of the library,
 * SDPV_SYMEIG_STATIC_WORK doubles (64 MiB) per GPU, held until the process ends; the call allocates
 * nothing. No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops). Sizes with 2 nb m^2 > SDPV_SYMEIG_STATIC_WORK (m = n
 * rounded up to even) return SDPV_ESIZE. Every call on one GPU shares its array, so calls that
 * overlap on one GPU (different streams, or host threads) race. */
#define SDPV_SYMEIG_STATIC_WORK (8LL << 20)
int sdpv_symeig_static(const double* A, int64_t nb, int64_t n, double* w, double* V, double* fro, double* off,
                       void* stream);

/* ---------------------------------------------------------------- SDPB binary BigFloat codec (host) */
/* n consecutive El::BigFloat records of boost's binary archive (uint8 is_zero; int32 size, int64 exp,
 * nl uint64 GMP limbs: value = sign(size) D 2^(64 (exp - |size|))) from byte pos of buf[len] into
 * MPF entries: out (n, L) int32 entry-major limbs holding the magnitude rounded half up to at most
 * 32L - 2 bits in two's complement, exps (n,) the binary exponents; zeros -> 0 and ez. Returns the
 * position after the last record, or -1 for a short or corrupt buffer. Host code only.
 * Replaces decode_bigfloats (sdpvibe/_bigfloat.c:16-59), with the same arguments. */
int64_t sdpv_decode_bigfloats(const uint8_t* buf, int64_t len, int64_t pos, int64_t n, int32_t nl, int32_t L,
                              int32_t* out, int64_t* exps, int64_t ez);

/* Inverse of sdpv_decode_bigfloats: n MPF entries (L limbs entry-major, exponents; zero when the
 * mantissa is 0 or the exponent is ez) -> records with nl limbs written to out[cap]; returns the bytes
 * written, or -1 when out is too small or a value needs more than nl limbs. Host code only.
 * Replaces encode_bigfloats (sdpvibe/_bigfloat.c:61-111), with the same arguments. */
int64_t sdpv_encode_bigfloats(const int32_t* in, const int64_t* exps, int64_t n, int32_t L, int32_t nl,
                              uint8_t* out, int64_t cap, int64_t ez);

/* ---------------------------------------------------------------- diagnostics */

/* D = A * B through one wave of v_mfma_i32_32x32x32_i8 using the same fragment mapping as the
 * GEMM kernels. A: (32, 32) int8 row-major, B: (32, 32) int8 row-major (k, n), D: (32, 32) int32.
 * No counterpart in the reference: a self-test of the matrix-core operand layout. */
int sdpv_mfma_selftest(const int8_t* A, const int8_t* B, int32_t* D, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SDPVIBE_HIP_H */
