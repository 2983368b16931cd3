/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY. A CPU restatement of the AntoineTilloy/SDPviBe
 * multiprecision semantics with GMP mpz integers, used by tests/, the smoke check and the CPU
 * baseline leg of bench.py as the checker of the HIP library. Nothing in sdpvibe_amd/ links,
 * loads or calls it.
 *
 * Every function works on host arrays in the library's layouts (limb planes: limb k of entry e at
 * X[k*plane + e], int32 two's complement; exponents int64; zero has e = EZ) and restates one
 * reference function, cited at its definition. The results are exact integers followed by the
 * reference's rounding rules; the algorithms (mpz arithmetic, direct CRT) are deliberately
 * different from the GPU ones (RNS on INT8 matrix cores).
 */
#include <gmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define EZ (-(1LL << 40))

typedef int64_t i64;
typedef uint32_t u32;

/* ---------------------------------------------------------------- mpz <-> limbs */
static void from_words(mpz_t x, const int32_t* X, i64 n, i64 stride) {
    u32* w = (u32*)malloc(sizeof(u32) * (size_t)n);
    for (i64 k = 0; k < n; k++) w[k] = (u32)X[k * stride];
    mpz_import(x, (size_t)n, -1, 4, 0, 0, w);
    if (w[n - 1] >> 31) {
        mpz_t t;
        mpz_init(t);
        mpz_ui_pow_ui(t, 2, (unsigned long)(32 * n));
        mpz_sub(x, x, t);
        mpz_clear(t);
    }
    free(w);
}

/* words 0..n-1 of x mod 2^(32n) */
static void to_words(const mpz_t x, int32_t* out, i64 n, i64 stride) {
    mpz_t r;
    mpz_init(r);
    mpz_fdiv_r_2exp(r, x, (mp_bitcnt_t)(32 * n));
    u32* w = (u32*)calloc((size_t)n + 2, sizeof(u32));
    size_t cnt = 0;
    mpz_export(w, &cnt, -1, 4, 0, 0, r);
    for (i64 k = 0; k < n; k++) out[k * stride] = (int32_t)w[k];
    free(w);
    mpz_clear(r);
}

static u32 word_of(const mpz_t x, i64 k) { /* word k of the infinite two's complement of x */
    mpz_t t;
    mpz_init(t);
    mpz_fdiv_q_2exp(t, x, (mp_bitcnt_t)(32 * k));
    mpz_fdiv_r_2exp(t, t, 32);
    u32 v = (u32)mpz_get_ui(t);
    mpz_clear(t);
    return v;
}

/* bit length of x, of ~x = -x-1 when x < 0 (limbops.py:142-151, 197-210) */
static i64 bitlen_tc(const mpz_t x) {
    if (mpz_sgn(x) >= 0) return mpz_sgn(x) == 0 ? 0 : (i64)mpz_sizeinbase(x, 2);
    mpz_t t;
    mpz_init(t);
    mpz_com(t, x);
    i64 r = mpz_sgn(t) == 0 ? 0 : (i64)mpz_sizeinbase(t, 2);
    mpz_clear(t);
    return r;
}

static i64 bitlen_abs(const mpz_t x) { return mpz_sgn(x) == 0 ? 0 : (i64)mpz_sizeinbase(x, 2); }

/* y = floor((x + 2^(s-1)) / 2^s) for s > 0 (half toward +inf), x * 2^-s for s <= 0.
 * |s| is clamped where the result no longer depends on it (see callers). */
static void round_shift(mpz_t y, const mpz_t x, i64 s, i64 clamp_right, i64 clamp_left) {
    if (s > 0) {
        if (s > clamp_right) { /* |x| < 2^(s-1): result 0 */
            mpz_set_ui(y, 0);
            return;
        }
        mpz_t h;
        mpz_init(h);
        mpz_setbit(h, (mp_bitcnt_t)(s - 1));
        mpz_add(y, x, h);
        mpz_fdiv_q_2exp(y, y, (mp_bitcnt_t)s);
        mpz_clear(h);
    } else {
        i64 t = -s;
        if (t > clamp_left) t = clamp_left; /* every kept word is 0 either way */
        mpz_mul_2exp(y, x, (mp_bitcnt_t)t);
    }
}

/* ---------------------------------------------------------------- limbops */

/* limbops.shift_round / _shift_round_kernel (sdpvibe/limbops.py:173-185, :32-62) */
void orc_shift_round(const int32_t* X, i64 Lin, i64 B, const int64_t* s, int32_t* out, i64 Lout,
                     int32_t* flag) {
    #pragma omp parallel
    {
        mpz_t x, y;
        mpz_init(x);
        mpz_init(y);
        #pragma omp for schedule(static)
        for (i64 e = 0; e < B; e++) {
            from_words(x, X + e, Lin, B);
            round_shift(y, x, s[e], 32 * Lin + 64, 32 * (Lout + 2));
            to_words(y, out + e, Lout, B);
            u32 top = word_of(y, Lout), prev = word_of(y, Lout - 1);
            u32 want = (prev >> 31) ? 0xffffffffu : 0u;
            if (flag) flag[e] = top != want;
        }
        mpz_clear(x);
        mpz_clear(y);
    }
}

/* limbops.add_same (sdpvibe/limbops.py:188-193, :66-81): A + sgn*B mod 2^(32L) */
void orc_add_same(const int32_t* A, const int32_t* Bm, int32_t* out, i64 L, i64 B, int32_t sgn) {
    mpz_t a, b;
    mpz_init(a);
    mpz_init(b);
    for (i64 e = 0; e < B; e++) {
        from_words(a, A + e, L, B);
        from_words(b, Bm + e, L, B);
        if (sgn < 0) mpz_sub(a, a, b); else mpz_add(a, a, b);
        to_words(a, out + e, L, B);
    }
    mpz_clear(a);
    mpz_clear(b);
}

/* limbops.bitlen_words (sdpvibe/limbops.py:213-218) */
void orc_bitlen(const int32_t* X, i64 L, i64 B, int64_t* out) {
    mpz_t x;
    mpz_init(x);
    for (i64 e = 0; e < B; e++) {
        from_words(x, X + e, L, B);
        out[e] = bitlen_tc(x);
    }
    mpz_clear(x);
}

/* batched.badd(..., 'align') / _badd_align_kernel (sdpvibe/batched.py:56-92) */
void orc_badd_align(const int32_t* A, const int32_t* Bm, const int64_t* EA, const int64_t* EB, int32_t* out,
                    int64_t* EO, i64 L, i64 nb, i64 rows, i64 cols, int32_t sgn) {
    i64 entries = nb * rows * cols;
    mpz_t a, b, ra, rb;
    mpz_init(a); mpz_init(b); mpz_init(ra); mpz_init(rb);
    for (i64 e = 0; e < entries; e++) {
        i64 col = (e / (rows * cols)) * cols + e % cols;
        i64 ex = (EA[col] > EB[col] ? EA[col] : EB[col]) + 1;
        from_words(a, A + e, L, entries);
        from_words(b, Bm + e, L, entries);
        round_shift(ra, a, ex - EA[col], 32 * L + 64, 0);
        round_shift(rb, b, ex - EB[col], 32 * L + 64, 0);
        if (sgn < 0) mpz_sub(ra, ra, rb); else mpz_add(ra, ra, rb);
        to_words(ra, out + e, L, entries);
        EO[col] = ex;
    }
    mpz_clear(a); mpz_clear(b); mpz_clear(ra); mpz_clear(rb);
}

/* ---------------------------------------------------------------- MPF (sdpvibe/mpf.py) */

/* mpf._top (sdpvibe/mpf.py:348-351) */
void orc_mpf_top(const int32_t* m, const int64_t* e, i64 L, i64 n, int64_t* top) {
    mpz_t x;
    mpz_init(x);
    for (i64 i = 0; i < n; i++) {
        if (e[i] == EZ) { top[i] = EZ; continue; }
        from_words(x, m + i, L, n);
        top[i] = e[i] + bitlen_tc(x);
    }
    mpz_clear(x);
}

/* Normalise the exact integer v (exponent ein) to L limbs: shift_mag of the magnitude
 * (half up), sign restored (mpf.py:36-56 with the mpf_norm / mpf_mul tails :84-88, :133-139). */
static void norm_value(const mpz_t v, i64 ein, i64 L, int32_t* om, i64 stride, int64_t* oe) {
    if (mpz_sgn(v) == 0) {
        for (i64 k = 0; k < L; k++) om[k * stride] = 0;
        *oe = EZ;
        return;
    }
    mpz_t mag, r;
    mpz_init(mag);
    mpz_init(r);
    mpz_abs(mag, v);
    i64 NB = 32 * L - 2;
    i64 s = bitlen_abs(mag) - NB;
    round_shift(r, mag, s, bitlen_abs(mag) + 64, 32 * (L + 2));
    mpz_fdiv_r_2exp(r, r, (mp_bitcnt_t)(32 * L)); /* words above L dropped (shift_mag nd = L) */
    if (mpz_sgn(v) < 0) mpz_neg(r, r);
    to_words(r, om, L, stride);
    *oe = ein + s;
    mpz_clear(mag);
    mpz_clear(r);
}

/* mpf.normalize / mpf_norm (sdpvibe/mpf.py:307-313, :127-140) */
void orc_mpf_norm(const int32_t* X, const int64_t* eI, i64 LI, int32_t* O, int64_t* eO, i64 L, i64 n) {
    mpz_t x;
    mpz_init(x);
    for (i64 i = 0; i < n; i++) {
        from_words(x, X + i, LI, n);
        norm_value(x, eI[i], L, O + i, n, eO + i);
    }
    mpz_clear(x);
}

/* mpf.mul / mpf_mul (sdpvibe/mpf.py:316-323, :58-89) */
void orc_mpf_mul(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB, int32_t* O,
                 int64_t* eO, i64 L, i64 n, int32_t bscalar) {
    mpz_t a, b;
    mpz_init(a);
    mpz_init(b);
    for (i64 i = 0; i < n; i++) {
        i64 ib = bscalar ? 0 : i, nb = bscalar ? 1 : n;
        from_words(a, A + i, L, n);
        from_words(b, Bm + ib, L, nb);
        mpz_mul(a, a, b);
        norm_value(a, eA[i] + eB[ib], L, O + i, n, eO + i);
    }
    mpz_clear(a);
    mpz_clear(b);
}

/* shift_mag(src, ns, s, dst, nd) of a magnitude, as an integer mod 2^(32 nd) (mpf.py:36-56) */
static void shift_mag_z(mpz_t y, const mpz_t mag, i64 s, i64 nd) {
    round_shift(y, mag, s, bitlen_abs(mag) + 64, 32 * (nd + 2));
    mpz_fdiv_r_2exp(y, y, (mp_bitcnt_t)(32 * nd));
}

/* mpf.add / mpf_add (sdpvibe/mpf.py:326-331, :92-124) */
void orc_mpf_add(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB, int32_t* O,
                 int64_t* eO, i64 L, i64 n, int32_t sgn) {
    mpz_t a, b, wa, wb, mod;
    mpz_init(a); mpz_init(b); mpz_init(wa); mpz_init(wb); mpz_init(mod);
    i64 W = L + 2;
    mpz_ui_pow_ui(mod, 2, (unsigned long)(32 * W));
    for (i64 i = 0; i < n; i++) {
        from_words(a, A + i, L, n);
        from_words(b, Bm + i, L, n);
        int na = mpz_sgn(a) < 0, nbg = (mpz_sgn(b) < 0) ^ (sgn < 0);
        mpz_abs(a, a);
        mpz_abs(b, b);
        i64 ea = eA[i], eb = eB[i];
        if (mpz_sgn(a) == 0) ea = EZ;
        if (mpz_sgn(b) == 0) eb = EZ;
        if (ea == EZ && eb == EZ) {
            for (i64 k = 0; k < L; k++) O[k * n + i] = 0;
            eO[i] = EZ;
            continue;
        }
        i64 t = (ea > eb ? ea : eb) - 32;
        i64 da = ea - t, db = eb - t, cl = 32 * (L + 3);
        i64 sa = da < -cl ? cl : -da, sb = db < -cl ? cl : -db;
        if (ea == EZ) mpz_set_ui(wa, 0); else shift_mag_z(wa, a, sa, W);
        if (eb == EZ) mpz_set_ui(wb, 0); else shift_mag_z(wb, b, sb, W);
        if (na) mpz_neg(wa, wa);
        if (nbg) mpz_neg(wb, wb);
        mpz_add(wa, wa, wb);
        mpz_fdiv_r(wa, wa, mod); /* W-word two's complement */
        if (mpz_tstbit(wa, (mp_bitcnt_t)(32 * W - 1))) mpz_sub(wa, wa, mod);
        norm_value(wa, t, L, O + i, n, eO + i);
    }
    mpz_clear(a); mpz_clear(b); mpz_clear(wa); mpz_clear(wb); mpz_clear(mod);
}

/* ---------------------------------------------------------------- RNS (sdpvibe/rns.py) */

/* RNS.residues (sdpvibe/rns.py:196-214): centred x mod m of the entries X[:, r, k0 + k] of an
 * (L, rows, K) limb-plane matrix; out (G, rows, kc) int32 = r in [-(m-1)/2, (m-1)/2]. */
void orc_residues(const int32_t* X, i64 L, i64 rows, i64 K, i64 k0, i64 kc, const int32_t* mods, i64 G,
                  int32_t* out) {
    #pragma omp parallel
    {
        mpz_t x;
        mpz_init(x);
        #pragma omp for schedule(static)
        for (i64 r = 0; r < rows; r++)
            for (i64 k = 0; k < kc; k++) {
                from_words(x, X + r * K + k0 + k, L, rows * K);
                for (i64 g = 0; g < G; g++) {
                    long m = mods[g];
                    long v = (long)mpz_fdiv_ui(x, (unsigned long)m);
                    if (v > m / 2) v -= m;
                    out[(g * rows + r) * kc + k] = (int32_t)v;
                }
            }
        mpz_clear(x);
    }
}

/* RNS.crt_block (sdpvibe/rns.py:227-245): the integer in (-M/2, M/2] congruent to acc[g*B + q]
 * modulo every m_g, as nw words W[t*B + q] and its bit length (of ~v for v < 0). Direct CRT
 * with mpz (sum of r_g * (M/m_g) * ((M/m_g)^-1 mod m_g), reduced mod M, centred). */
void orc_crt(const int16_t* acc, const int32_t* mods, i64 n, i64 B, int32_t* W, i64 nw, int64_t* bitlen) {
    mpz_t M, Mi, c, v, half;
    mpz_init_set_ui(M, 1);
    mpz_init(Mi); mpz_init(c); mpz_init(v); mpz_init(half);
    for (i64 g = 0; g < n; g++) mpz_mul_ui(M, M, (unsigned long)mods[g]);
    mpz_t* coef = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)n);
    for (i64 g = 0; g < n; g++) {
        mpz_init(coef[g]);
        mpz_divexact_ui(Mi, M, (unsigned long)mods[g]);
        mpz_set_ui(c, (unsigned long)mods[g]);
        mpz_invert(coef[g], Mi, c);
        mpz_mul(coef[g], coef[g], Mi);
    }
    mpz_fdiv_q_2exp(half, M, 1);
    for (i64 q = 0; q < B; q++) {
        mpz_set_ui(v, 0);
        for (i64 g = 0; g < n; g++) {
            long r = acc[g * B + q] % mods[g];
            if (r < 0) r += mods[g];
            mpz_addmul_ui(v, coef[g], (unsigned long)r);
        }
        mpz_fdiv_r(v, v, M);
        if (mpz_cmp(v, half) > 0) mpz_sub(v, v, M);
        to_words(v, W + q, nw, B);
        if (bitlen) bitlen[q] = bitlen_tc(v);
    }
    for (i64 g = 0; g < n; g++) mpz_clear(coef[g]);
    free(coef);
    mpz_clear(M); mpz_clear(Mi); mpz_clear(c); mpz_clear(v); mpz_clear(half);
}

/* ---------------------------------------------------------------- GEMM */

/* Operand geometry: entry (b, r, k) at b*ldb + r*ldr + k*ldk of every limb plane and of e. */
typedef struct { i64 nb, rows, K, ldb, ldr, ldk, plane; } geom_t;

/* mpf._align_rows (sdpvibe/mpf.py:390-398): per row, E = max top - NB (0 for a zero row) and the
 * entries shift-rounded to L limbs at 2^E, returned as exact integers a[(b*rows + r)*K + k]. */
static void align_rows(const int32_t* m, const int64_t* e, i64 L, geom_t g, mpz_t* a, int64_t* Erow) {
    i64 NB = 32 * L - 2;
    #pragma omp parallel
    {
        mpz_t x, y;
        mpz_init(x);
        mpz_init(y);
        int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);
        #pragma omp for schedule(static)
        for (i64 R = 0; R < g.nb * g.rows; R++) {
            i64 b = R / g.rows, r = R % g.rows, base = b * g.ldb + r * g.ldr;
            i64 mx = EZ;
            for (i64 k = 0; k < g.K; k++) {
                i64 off = base + k * g.ldk;
                if (e[off] == EZ) continue;
                from_words(x, m + off, L, g.plane);
                i64 t = e[off] + bitlen_tc(x);
                if (t > mx) mx = t;
            }
            i64 E = mx - NB;
            if (E < EZ / 2) E = 0;
            Erow[R] = E;
            for (i64 k = 0; k < g.K; k++) {
                i64 off = base + k * g.ldk;
                i64 s = (e[off] == EZ) ? (1LL << 30) : E - e[off];
                from_words(x, m + off, L, g.plane);
                round_shift(y, x, s, 32 * L + 64, 32 * (L + 2));
                /* wrap to L words (two's complement), as the L-limb shift_round output */
                to_words(y, tmp, L, 1);
                from_words(a[R * g.K + k], tmp, L, 1);
            }
        }
        free(tmp);
        mpz_clear(x);
        mpz_clear(y);
    }
}

/* mpf.gemm (sdpvibe/mpf.py:401-452): C[b] = op(A)[b] op(B)[b], rows of op(A) and columns of op(B)
 * aligned (first rounding), exact products, then normalised per entry (second rounding).
 * ga describes op(A) (rows M), gb describes op(B)^T (rows N); out (L, nb, M, N), oe (nb, M, N). */
void orc_mpf_gemm(const int32_t* Am, const int64_t* Ae, i64 L, i64 nb, i64 M, i64 N, i64 K, i64 aldb, i64 aldr,
                  i64 aldk, i64 aplane, const int32_t* Bm, const int64_t* Be, i64 bldb, i64 bldr, i64 bldk,
                  i64 bplane, int32_t* out, int64_t* oe) {
    geom_t ga = {nb, M, K, aldb, aldr, aldk, aplane}, gb = {nb, N, K, bldb, bldr, bldk, bplane};
    mpz_t* a = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)(nb * M * K));
    mpz_t* b = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)(nb * N * K));
    for (i64 i = 0; i < nb * M * K; i++) mpz_init(a[i]);
    for (i64 i = 0; i < nb * N * K; i++) mpz_init(b[i]);
    int64_t* EA = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb * M));
    int64_t* EB = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb * N));
    align_rows(Am, Ae, L, ga, a, EA);
    align_rows(Bm, Be, L, gb, b, EB);
    i64 plane = nb * M * N;
    #pragma omp parallel
    {
        mpz_t c;
        mpz_init(c);
        #pragma omp for schedule(dynamic, 16)
        for (i64 q = 0; q < nb * M * N; q++) {
            i64 bb = q / (M * N), i = (q / N) % M, j = q % N;
            mpz_set_ui(c, 0);
            const mpz_t* ar = a + (bb * M + i) * K;
            const mpz_t* br = b + (bb * N + j) * K;
            for (i64 k = 0; k < K; k++) mpz_addmul(c, ar[k], br[k]);
            norm_value(c, EA[bb * M + i] + EB[bb * N + j], L, out + q, plane, oe + q);
        }
        mpz_clear(c);
    }
    for (i64 i = 0; i < nb * M * K; i++) mpz_clear(a[i]);
    for (i64 i = 0; i < nb * N * K; i++) mpz_clear(b[i]);
    free(a); free(b); free(EA); free(EB);
}

/* Exact product of fixed-point limb matrices: C = A Bt^T with A (M, K), Bt (N, K) limb planes
 * (L limbs), as nw words per entry (W[t*M*N + i*N + j]); the exact core of mpmat.gemm_nt
 * (sdpvibe/mpmat.py:169-202) before its rounding. */
void orc_gemm_exact(const int32_t* A, const int32_t* Bt, i64 L, i64 M, i64 N, i64 K, int32_t* W, i64 nw) {
    mpz_t* a = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)(M * K));
    mpz_t* b = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)(N * K));
    for (i64 i = 0; i < M * K; i++) { mpz_init(a[i]); from_words(a[i], A + i, L, M * K); }
    for (i64 i = 0; i < N * K; i++) { mpz_init(b[i]); from_words(b[i], Bt + i, L, N * K); }
    #pragma omp parallel
    {
        mpz_t c;
        mpz_init(c);
        #pragma omp for schedule(dynamic, 16)
        for (i64 q = 0; q < M * N; q++) {
            i64 i = q / N, j = q % N;
            mpz_set_ui(c, 0);
            for (i64 k = 0; k < K; k++) mpz_addmul(c, a[i * K + k], b[j * K + k]);
            to_words(c, W + q, nw, M * N);
        }
        mpz_clear(c);
    }
    for (i64 i = 0; i < M * K; i++) mpz_clear(a[i]);
    for (i64 i = 0; i < N * K; i++) mpz_clear(b[i]);
    free(a); free(b);
}

/* batched.bgemm_nt + _bcrt_round (sdpvibe/batched.py:165-231): C[b] = A[b] Bt[b]^T exactly (A: (L, nb, M, K),
 * Bt: (L, nb, N, K) fixed-point limb planes), then rounded to L limbs per (block, column):
 *   mode 0 (fixed): exponent oexp[b*N + j] given; shift = oexp - e_in
 *   mode 1 ('col'): shift = max_i bitlen(C[b, i, j]) - (32L - 32); oexp = e_in + shift
 *   mode 2 ('mat'): 'col', then every column of block b re-rounded to the block's max exponent
 * e_in (nb, N) is the exponent of the exact product. out (L, nb, M, N), oexp (nb, N) in/out. */
void orc_bgemm_nt(const int32_t* A, const int32_t* Bt, i64 L, i64 nb, i64 M, i64 N, i64 K, const int64_t* e_in,
                  int32_t mode, int32_t* out, int64_t* oexp) {
    mpz_t* c = (mpz_t*)malloc(sizeof(mpz_t) * (size_t)(nb * M * N));
    i64 plA = nb * M * K, plB = nb * N * K, plC = nb * M * N;
    #pragma omp parallel
    {
        mpz_t a, b;
        mpz_init(a);
        mpz_init(b);
        #pragma omp for schedule(dynamic, 8)
        for (i64 q = 0; q < plC; q++) {
            i64 bb = q / (M * N), i = (q / N) % M, j = q % N;
            mpz_init_set_ui(c[q], 0);
            for (i64 k = 0; k < K; k++) {
                from_words(a, A + (bb * M + i) * K + k, L, plA);
                from_words(b, Bt + (bb * N + j) * K + k, L, plB);
                mpz_addmul(c[q], a, b);
            }
        }
        mpz_clear(a);
        mpz_clear(b);
    }
    i64 H = 32 * L - 32;
    for (i64 bb = 0; bb < nb; bb++)
        for (i64 j = 0; j < N; j++) {
            i64 ej = bb * N + j;
            if (mode != 0) {
                i64 mx = 0;
                for (i64 i = 0; i < M; i++) {
                    i64 bl = bitlen_tc(c[(bb * M + i) * N + j]);
                    if (i == 0 || bl > mx) mx = bl;
                }
                oexp[ej] = e_in[ej] + (mx - H);
            }
            mpz_t y;
            mpz_init(y);
            for (i64 i = 0; i < M; i++) {
                i64 q = (bb * M + i) * N + j;
                round_shift(y, c[q], oexp[ej] - e_in[ej], 1LL << 40, 32 * (L + 2));
                to_words(y, out + q, L, plC);
            }
            mpz_clear(y);
        }
    if (mode == 2) {
        mpz_t x, y;
        mpz_init(x);
        mpz_init(y);
        for (i64 bb = 0; bb < nb; bb++) {
            i64 emax = oexp[bb * N];
            for (i64 j = 1; j < N; j++) if (oexp[bb * N + j] > emax) emax = oexp[bb * N + j];
            for (i64 j = 0; j < N; j++) {
                for (i64 i = 0; i < M; i++) {
                    i64 q = (bb * M + i) * N + j;
                    from_words(x, out + q, L, plC);
                    round_shift(y, x, emax - oexp[bb * N + j], 32 * L + 64, 32 * (L + 2));
                    to_words(y, out + q, L, plC);
                }
                oexp[bb * N + j] = emax;
            }
        }
        mpz_clear(x);
        mpz_clear(y);
    }
    for (i64 q = 0; q < plC; q++) mpz_clear(c[q]);
    free(c);
}

/* ---------------------------------------------------------------- tile base case (sdpvibe/basecase.py)
 * Wide fixed point: WL 32-bit words, two's complement, F = 32*LF fractional bits. Restated from the
 * CUDA source held in basecase.py:13-246 (read as text). */
typedef struct { int WL, LF; } wfx_t;

static int w_isneg(const u32* a, int WL) { return a[WL - 1] >> 31; }
static void w_negate(u32* a, int WL) {
    uint64_t c = 1;
    for (int i = 0; i < WL; i++) { uint64_t t = (uint64_t)(~a[i]) + c; a[i] = (u32)t; c = t >> 32; }
}
static void w_addto(u32* r, const u32* a, int WL) {
    uint64_t c = 0;
    for (int i = 0; i < WL; i++) { uint64_t t = (uint64_t)r[i] + a[i] + c; r[i] = (u32)t; c = t >> 32; }
}
static void w_subfrom(u32* r, const u32* a, int WL) {
    int64_t br = 0;
    for (int i = 0; i < WL; i++) { int64_t t = (int64_t)r[i] - (int64_t)a[i] + br; r[i] = (u32)t; br = t >> 32; }
}
static int w_topl(const u32* a, int WL) { for (int i = WL - 1; i > 0; i--) if (a[i]) return i; return 0; }
static void w_shr1(u32* a, int WL) {
    for (int i = 0; i < WL - 1; i++) a[i] = (a[i] >> 1) | (a[i + 1] << 31);
    a[WL - 1] = (u32)(((int32_t)a[WL - 1]) >> 1);
}

/* r = (a*b) >> F, signed, truncated; t > 0: only the top t words of |a| and |b| (basecase.py:31-53) */
static void w_mulfx(u32* r, const u32* a, const u32* b, int t, wfx_t f) {
    int WL = f.WL, LF = f.LF;
    u32 x[256], y[256];
    memcpy(x, a, 4 * (size_t)WL);
    memcpy(y, b, 4 * (size_t)WL);
    int sa = w_isneg(x, WL), sb = w_isneg(y, WL);
    if (sa) w_negate(x, WL);
    if (sb) w_negate(y, WL);
    int ta = w_topl(x, WL), tb = w_topl(y, WL);
    int la = t ? (ta - t + 1 > 0 ? ta - t + 1 : 0) : 0, lb = t ? (tb - t + 1 > 0 ? tb - t + 1 : 0) : 0;
    int kmin = LF - 2 > la + lb ? LF - 2 : la + lb, kmax = ta + tb;
    for (int j = 0; j < WL; j++) r[j] = 0;
    uint64_t carry = 0, hprev = 0;
    for (int k = kmin; k <= kmax + 2 && k < LF + WL; k++) {
        uint64_t lo = 0, hi = 0;
        if (k <= kmax) {
            int i0 = la > k - tb ? la : k - tb, i1 = ta < k - lb ? ta : k - lb;
            for (int i = i0; i <= i1; i++) { uint64_t p = (uint64_t)x[i] * y[k - i]; lo += (u32)p; hi += p >> 32; }
        }
        uint64_t s = lo + hprev + carry;
        if (k >= LF) r[k - LF] = (u32)s;
        carry = s >> 32;
        hprev = hi;
    }
    if (sa ^ sb) w_negate(r, WL);
}

static double w_to_double(const u32* a, wfx_t f) {                 /* basecase.py:59-68 */
    int WL = f.WL;
    u32 x[256];
    memcpy(x, a, 4 * (size_t)WL);
    int s = w_isneg(x, WL);
    if (s) w_negate(x, WL);
    int t = w_topl(x, WL);
    double v = (double)x[t];
    if (t >= 1) v = v * 4294967296.0 + (double)x[t - 1];
    if (t >= 2) v = v * 4294967296.0 + (double)x[t - 2];
    int e = 32 * (t - 2 > 0 ? t - 2 : 0) - 32 * f.LF;
    v = ldexp(v, e);
    return s ? -v : v;
}

static void w_from_double(u32* r, double v, wfx_t f) {             /* basecase.py:69-78, v > 0 */
    int WL = f.WL;
    for (int i = 0; i < WL; i++) r[i] = 0;
    if (!(v > 0.0) || !isfinite(v)) return;   /* non-positive pivot: reciprocal 0 (undefined in the reference) */
    int e;
    double m = frexp(v, &e);
    uint64_t mi = (uint64_t)ldexp(m, 53);
    int pos = e - 53 + 32 * f.LF;
    if (pos < 0) { mi = (-pos < 64) ? (mi >> (-pos)) : 0; pos = 0; }
    int li = pos / 32, sh = pos % 32;
    unsigned __int128 w = ((unsigned __int128)mi) << sh;
    for (int k = 0; k < 4 && li + k < WL; k++) { r[li + k] = (u32)(w & 0xFFFFFFFFu); w >>= 32; }
}

/* warp-cooperative d^{-1/2} of chol_tiles (basecase.py:145-156): same products as mulfx */
static void w_rsqrt(u32* r, const u32* d, wfx_t f) {
    int WL = f.WL;
    u32 t1[256], t2[256], one[256];
    w_from_double(r, 1.0 / sqrt(w_to_double(d, f)), f);
    for (int i = 0; i < WL; i++) one[i] = 0;
    one[f.LF] = 1;
    for (int bits = 100;; bits *= 2) {
        int t = bits / 32 + 2;
        if (t >= WL) t = 0;
        w_mulfx(t1, r, r, t, f);
        w_mulfx(t2, t1, d, t, f);
        memcpy(t1, one, 4 * (size_t)WL);
        w_subfrom(t1, t2, WL);
        w_mulfx(t2, r, t1, t, f);
        w_shr1(t2, WL);
        w_addto(r, t2, WL);
        if (t == 0) break;
    }
}

/* chol_tiles + upper-triangle clearing + trinv_tiles (basecase.py:159-226, batched.py:250-255) on nb
 * tiles of BT x BT wide entries. W: (nb, BT, BT, WL) entry-major words, factored in place (lower L,
 * upper cleared); RINV (nb, BT, WL) pivot reciprocals; X (nb, BT, BT, WL) = L^-1. */
void orc_tile_chol(uint32_t* W, uint32_t* RINV, uint32_t* X, i64 nb, int BT, int WL, int LF) {
    wfx_t f = {WL, LF};
    #pragma omp parallel for schedule(dynamic, 1)
    for (i64 b = 0; b < nb; b++) {
        u32* T = W + (size_t)b * BT * BT * WL;
        u32* RI = RINV + (size_t)b * BT * WL;
        u32* XT = X + (size_t)b * BT * BT * WL;
        u32 v[256];
        for (int k = 0; k < BT; k++) {
            u32* akk = T + ((size_t)k * BT + k) * WL;
            u32 r[256];
            w_rsqrt(r, akk, f);
            w_mulfx(v, akk, r, 0, f);
            memcpy(akk, v, 4 * (size_t)WL);
            memcpy(RI + (size_t)k * WL, r, 4 * (size_t)WL);
            for (int i = k + 1; i < BT; i++) {
                u32* aik = T + ((size_t)i * BT + k) * WL;
                w_mulfx(v, aik, RI + (size_t)k * WL, 0, f);
                memcpy(aik, v, 4 * (size_t)WL);
            }
            for (int i = k + 1; i < BT; i++)
                for (int j = k + 1; j <= i; j++) {
                    w_mulfx(v, T + ((size_t)i * BT + k) * WL, T + ((size_t)j * BT + k) * WL, 0, f);
                    w_subfrom(T + ((size_t)i * BT + j) * WL, v, WL);
                }
        }
        for (int i = 0; i < BT; i++)
            for (int j = i + 1; j < BT; j++) memset(T + ((size_t)i * BT + j) * WL, 0, 4 * (size_t)WL);
        memset(XT, 0, 4 * (size_t)BT * BT * WL);
        for (int i = 0; i < BT; i++) {
            for (int k = 0; k <= i; k++) {
                u32* xik = XT + ((size_t)i * BT + k) * WL;
                if (k == i) { memcpy(xik, RI + (size_t)i * WL, 4 * (size_t)WL); continue; }
                u32 acc[256];
                memset(acc, 0, 4 * (size_t)WL);
                for (int j = k; j < i; j++) {
                    w_mulfx(v, T + ((size_t)i * BT + j) * WL, XT + ((size_t)j * BT + k) * WL, 0, f);
                    w_addto(acc, v, WL);
                }
                w_mulfx(v, acc, RI + (size_t)i * WL, 0, f);
                w_negate(v, WL);
                memcpy(xik, v, 4 * (size_t)WL);
            }
        }
    }
}

/* trsv_warp (sdpvibe/basecase.py:228-245): L y = v in place for nvec vectors V (nvec, BT, WL) with the
 * factored tile Lw (BT, BT, WL) and pivot reciprocals RINV (BT, WL); y_j = mulfx(v_j, rinv_j), then
 * v_i -= mulfx(l_ij, y_j) for i > j. */
void orc_tile_trsv(const uint32_t* Lw, const uint32_t* RINV, uint32_t* V, i64 nvec, int BT, int WL, int LF) {
    wfx_t f = {WL, LF};
    #pragma omp parallel for schedule(dynamic, 4)
    for (i64 q = 0; q < nvec; q++) {
        u32* vb = V + (size_t)q * BT * WL;
        u32 t[256];
        for (int j = 0; j < BT; j++) {
            u32* vj = vb + (size_t)j * WL;
            w_mulfx(t, vj, RINV + (size_t)j * WL, 0, f);
            memcpy(vj, t, 4 * (size_t)WL);
            for (int i = j + 1; i < BT; i++) {
                w_mulfx(t, Lw + ((size_t)i * BT + j) * WL, vj, 0, f);
                w_subfrom(vb + (size_t)i * WL, t, WL);
            }
        }
    }
}
