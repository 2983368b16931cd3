// Limb-level bookkeeping and per-entry multiprecision floats on gfx950.
// One thread per entry; words are streamed straight from the limb planes (coalesced across
// entries) or kept in per-thread LDS slices, so a single kernel serves every limb count.
#include "common.hpp"

namespace {

struct PlaneSrc {  // limbs of one entry in limb-plane layout
    const i32* p;
    i64 stride;
    __device__ __forceinline__ u32 operator()(int k) const { return (u32)p[(i64)k * stride]; }
};

struct LdsSrc {  // per-thread words in LDS, thread-strided (bank-conflict free)
    const u32* p;
    int stride;
    __device__ __forceinline__ u32 operator()(int k) const { return p[k * stride]; }
};

constexpr int TPB = 256;

// ------------------------------------------------------------------ shift_round
__global__ void __launch_bounds__(TPB) k_shift_round(const i32* __restrict__ X, i64 B, int Lin,
                                                     const i64* __restrict__ S, i32* __restrict__ O,
                                                     int Lout, i32* __restrict__ flag, int per_entry) {
    const i64 e = blockIdx.x * (i64)TPB + threadIdx.x;
    if (e >= B) return;
    PlaneSrc src{X + e, B};
    const int bad = shr_round(src, Lin, S[e], Lout, [&](int k, u32 w) { O[(i64)k * B + e] = (i32)w; });
    if (per_entry) {
        flag[e] = bad;
    } else if (bad) {
        atomicOr(flag, 1);
    }
}

// ------------------------------------------------------------------ add_same
__global__ void __launch_bounds__(TPB) k_add_same(const i32* __restrict__ A, const i32* __restrict__ Bm,
                                                  i32* __restrict__ O, int L, i64 B, int sub) {
    const i64 e = blockIdx.x * (i64)TPB + threadIdx.x;
    if (e >= B) return;
    u64 c = sub ? 1u : 0u;
    for (int k = 0; k < L; k++) {
        const u32 a = (u32)A[(i64)k * B + e];
        u32 b = (u32)Bm[(i64)k * B + e];
        if (sub) b = ~b;
        const u64 t = (u64)a + b + c;
        c = t >> 32;
        O[(i64)k * B + e] = (i32)(u32)t;
    }
}

// ------------------------------------------------------------------ bitlen / top
__global__ void __launch_bounds__(TPB) k_bitlen(const i32* __restrict__ X, int L, i64 B, i64* __restrict__ out) {
    const i64 e = blockIdx.x * (i64)TPB + threadIdx.x;
    if (e >= B) return;
    out[e] = bitlen_tc(PlaneSrc{X + e, B}, L);
}

__global__ void __launch_bounds__(TPB) k_mpf_top(const i32* __restrict__ m, const i64* __restrict__ ev, int L,
                                                 i64 n, i64* __restrict__ top) {
    const i64 e = blockIdx.x * (i64)TPB + threadIdx.x;
    if (e >= n) return;
    const i64 ex = ev[e];
    top[e] = (ex == EZV) ? EZV : ex + bitlen_tc(PlaneSrc{m + e, n}, L);
}

// ------------------------------------------------------------------ batched aligned add
// batched.py:56-79: both operands shift right by >= 1 bit (round half up), then add (wraps).
__global__ void __launch_bounds__(TPB) k_badd_align(const i32* __restrict__ A, const i32* __restrict__ Bm,
                                                    const i64* __restrict__ EA, const i64* __restrict__ EB,
                                                    i32* __restrict__ O, i64* __restrict__ EO, int L,
                                                    i64 entries, i64 rows, i64 cols, int sub) {
    const i64 e = blockIdx.x * (i64)TPB + threadIdx.x;
    if (e >= entries) return;
    const i64 col = (e / (rows * cols)) * cols + e % cols;
    const i64 ea = EA[col], eb = EB[col];
    const i64 ex = (ea > eb ? ea : eb) + 1;
    PlaneSrc sa{A + e, entries}, sb{Bm + e, entries};
    // Stream both rounded operands word by word and add them with a running carry.
    // Each rounded operand fits L words (a right shift of at least one bit).
    const u32 exta = (sa(L - 1) >> 31) ? 0xffffffffu : 0u;
    const u32 extb = (sb(L - 1) >> 31) ? 0xffffffffu : 0u;
    const i64 s_a = ex - ea, s_b = ex - eb;
    const i64 qa = s_a >> 5, qb = s_b >> 5;
    const int ra = (int)(s_a & 31), rb = (int)(s_b & 31);
    u64 ca = (ext_word(sa, L, exta, (s_a - 1) >> 5) >> (int)((s_a - 1) & 31)) & 1u;
    u64 cb = (ext_word(sb, L, extb, (s_b - 1) >> 5) >> (int)((s_b - 1) & 31)) & 1u;
    u64 c = sub ? 1u : 0u;
    for (int k = 0; k < L; k++) {
        const u64 a = ((((u64)ext_word(sa, L, exta, qa + k)) | ((u64)ext_word(sa, L, exta, qa + k + 1) << 32)) >> ra & 0xffffffffull) + ca;
        const u64 b = ((((u64)ext_word(sb, L, extb, qb + k)) | ((u64)ext_word(sb, L, extb, qb + k + 1) << 32)) >> rb & 0xffffffffull) + cb;
        ca = a >> 32;
        cb = b >> 32;
        u32 bw = (u32)b;
        if (sub) bw = ~bw;
        const u64 t = (u64)(u32)a + bw + c;
        c = t >> 32;
        O[(i64)k * entries + e] = (i32)(u32)t;
    }
    if ((e / cols) % rows == 0) EO[col] = ex;
}

// ------------------------------------------------------------------ mpf_norm (mpf.py:127-140)
__global__ void __launch_bounds__(TPB) k_mpf_norm(const i32* __restrict__ X, const i64* __restrict__ eI, int LI,
                                                  i32* __restrict__ O, i64* __restrict__ eO, int L, i64 n) {
    const i64 i = blockIdx.x * (i64)TPB + threadIdx.x;
    if (i >= n) return;
    const int NB = 32 * L - 2;
    PlaneSrc x{X + i, n};
    MagView<PlaneSrc> mag(x, LI);
    const int bl = bitlen_mag(mag, LI);
    if (bl == 0) {
        for (int k = 0; k < L; k++) O[(i64)k * n + i] = 0;
        eO[i] = EZV;
        return;
    }
    const int s = bl - NB;
    NegStream ng(mag.neg);
    shift_mag(mag, LI, s, L, [&](int k, u32 w) { O[(i64)k * n + i] = (i32)ng(w); });
    eO[i] = eI[i] + s;
}

// ------------------------------------------------------------------ mpf_mul (mpf.py:58-89)
// Magnitudes staged in per-thread LDS slices; product scanning over runtime L.
constexpr int MUL_TPB = 64;
__global__ void __launch_bounds__(MUL_TPB) k_mpf_mul(const i32* __restrict__ A, const i64* __restrict__ eA,
                                                     const i32* __restrict__ Bm, const i64* __restrict__ eB,
                                                     i32* __restrict__ O, i64* __restrict__ eO, int L, i64 n,
                                                     int bscalar) {
    SDPV_DYNAMIC_LDS(u32, lds);
    const int t = threadIdx.x;
    const i64 i = blockIdx.x * (i64)MUL_TPB + t;
    if (i >= n) return;
    const i64 ib = bscalar ? 0 : i, nb = bscalar ? 1 : n;
    u32* a = lds + t;                    // a[k * MUL_TPB]
    u32* b = lds + L * MUL_TPB + t;      // b[k * MUL_TPB]
    u32* p = lds + 2 * L * MUL_TPB + t;  // p[k * MUL_TPB], 2L words
    PlaneSrc xa{A + i, n}, xb{Bm + ib, nb};
    MagView<PlaneSrc> ma(xa, L), mb(xb, L);
    for (int k = 0; k < L; k++) {
        a[k * MUL_TPB] = ma(k);
        b[k * MUL_TPB] = mb(k);
    }
    LdsSrc la{a, MUL_TPB}, lb{b, MUL_TPB};
    if (bitlen_mag(la, L) == 0 || bitlen_mag(lb, L) == 0) {
        for (int k = 0; k < L; k++) O[(i64)k * n + i] = 0;
        eO[i] = EZV;
        return;
    }
    u64 carry = 0, hprev = 0;
    for (int k = 0; k < 2 * L - 1; k++) {
        u64 lo = 0, hi = 0;
        const int first = k - L + 1 > 0 ? k - L + 1 : 0, last = k < L - 1 ? k : L - 1;
        for (int x = first; x <= last; x++) {
            const u64 pr = (u64)a[x * MUL_TPB] * b[(k - x) * MUL_TPB];
            lo += (u32)pr;
            hi += pr >> 32;
        }
        const u64 sum = lo + hprev + carry;
        p[k * MUL_TPB] = (u32)sum;
        carry = sum >> 32;
        hprev = hi;
    }
    p[(2 * L - 1) * MUL_TPB] = (u32)(hprev + carry);
    LdsSrc lp{p, MUL_TPB};
    const int NB = 32 * L - 2;
    const int s = bitlen_mag(lp, 2 * L) - NB;
    NegStream ng(ma.neg != mb.neg);
    shift_mag(lp, 2 * L, s, L, [&](int k, u32 w) { O[(i64)k * n + i] = (i32)ng(w); });
    eO[i] = eA[i] + eB[ib] + s;
}

// ------------------------------------------------------------------ mpf_add (mpf.py:92-124)
constexpr int ADD_TPB = 64;
__global__ void __launch_bounds__(ADD_TPB) k_mpf_add(const i32* __restrict__ A, const i64* __restrict__ eA,
                                                     const i32* __restrict__ Bm, const i64* __restrict__ eB,
                                                     i32* __restrict__ O, i64* __restrict__ eO, int L, i64 n,
                                                     int sgn) {
    SDPV_DYNAMIC_LDS(u32, lds);
    const int t = threadIdx.x;
    const i64 i = blockIdx.x * (i64)ADD_TPB + t;
    if (i >= n) return;
    const int W = L + 2;
    u32* wa = lds + t;            // wa[k * ADD_TPB], W words
    PlaneSrc xa{A + i, n}, xb{Bm + i, n};
    MagView<PlaneSrc> ma(xa, L), mb(xb, L);
    const bool na = ma.neg, nbg = mb.neg ^ (sgn < 0);
    i64 ea = eA[i], eb = eB[i];
    const int la = bitlen_mag(ma, L), lb = bitlen_mag(mb, L);
    if (la == 0) ea = EZV;
    if (lb == 0) eb = EZV;
    if (ea == EZV && eb == EZV) {
        for (int k = 0; k < L; k++) O[(i64)k * n + i] = 0;
        eO[i] = EZV;
        return;
    }
    const i64 tt = (ea > eb ? ea : eb) - 32;
    const i64 da = ea - tt, db = eb - tt;
    const i64 clampv = 32 * (i64)(L + 3);
    const i64 sa = da < -clampv ? clampv : -da;
    const i64 sb = db < -clampv ? clampv : -db;
    // wa = +-round(|a| / 2^sa) in W words
    {
        NegStream ng(na);
        if (ea == EZV) {
            for (int k = 0; k < W; k++) wa[k * ADD_TPB] = 0;
        } else {
            shift_mag(ma, L, sa, W, [&](int k, u32 w) { wa[k * ADD_TPB] = ng(w); });
        }
    }
    // wa += +-round(|b| / 2^sb), streamed
    if (eb != EZV) {
        NegStream ng(nbg);
        u64 c = 0;
        shift_mag(mb, L, sb, W, [&](int k, u32 w) {
            const u64 v = (u64)wa[k * ADD_TPB] + ng(w) + c;
            wa[k * ADD_TPB] = (u32)v;
            c = v >> 32;
        });
    }
    LdsSrc sw{wa, ADD_TPB};
    MagView<LdsSrc> mw(sw, W);
    const int bl = bitlen_mag(mw, W);
    if (bl == 0) {
        for (int k = 0; k < L; k++) O[(i64)k * n + i] = 0;
        eO[i] = EZV;
        return;
    }
    const int NB = 32 * L - 2;
    const int s = bl - NB;
    NegStream ng(mw.neg);
    shift_mag(mw, W, s, L, [&](int k, u32 w) { O[(i64)k * n + i] = (i32)ng(w); });
    eO[i] = tt + s;
}

// ------------------------------------------------------------------ row alignment exponents
// Erow = max_k top(entry) - NB over entries with e != EZ (mpf.py:390-395), computed as a K-split
// reduction: a fill pass, then blocks over (rows, k slices) combine with 64-bit atomicMax.
// Rows without such entries keep SDPV_EZ (the reference then uses exponent 0; no product
// depends on it because every entry of such a row is zero).
__device__ __forceinline__ i64 entry_top(const i32* X, const i64* e, int L, i64 plane, i64 off) {
    const i64 ex = e[off];
    if (ex == EZV) return EZV;
    return ex + bitlen_tc(PlaneSrc{X + off, plane}, L);
}

// top of an entry when its exponent is known and its top limb is neither 0 nor -1 (normalised
// MPF); otherwise the full scan
__device__ __forceinline__ i64 entry_top_fast(const i32* X, i64 plane, int L, i64 off, i64 ex, u32 topw) {
    if (ex == EZV) return EZV;
    if (topw != 0u && topw != 0xffffffffu) {
        const u32 x = topw ^ ((topw >> 31) ? 0xffffffffu : 0u);
        return ex + 32 * (i64)(L - 1) + 32 - clz32(x);
    }
    return ex + bitlen_tc(PlaneSrc{X + off, plane}, L);
}

__device__ __forceinline__ i64 row_exp_of(i64 mx, int L) { return mx == EZV ? EZV : mx - (32 * (i64)L - 2); }

// rows contiguous in k (ldk == 1): one wave per row, 8 entries per lane in flight
__global__ void __launch_bounds__(256) k_row_exp_wave(const i32* __restrict__ X, const i64* __restrict__ e, int L,
                                                      i64 plane, i64 nrows, i64 rows, i64 K, i64 ldb, i64 ldr,
                                                      i64* __restrict__ Erow) {
    const i64 R = blockIdx.x * (i64)4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (R >= nrows) return;
    const i64 b = R / rows, r = R % rows;
    const i64 base = b * ldb + r * ldr;
    const i32* top = X + (i64)(L - 1) * plane;
    i64 mx = EZV;
    for (i64 k0 = 0; k0 < K; k0 += 8 * 64) {
        i64 ex[8];
        u32 tw[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const i64 k = k0 + u * 64 + lane;
            ex[u] = k < K ? e[base + k] : EZV;
            tw[u] = k < K ? (u32)top[base + k] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const i64 t = entry_top_fast(X, plane, L, base + k0 + u * 64 + lane, ex[u], tw[u]);
            mx = t > mx ? t : mx;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const i64 v = __shfl_xor(mx, o, 64);
        mx = v > mx ? v : mx;
    }
    if (lane == 0) Erow[R] = row_exp_of(mx, L);
}

// other layouts: 64 rows (lanes) per block of 16 waves, wave w takes k = w, w + 16, ... of the
// block's k range (blockIdx.y: one of gridDim.y parts of kchunk); the 16 partial maxima per row
// meet in LDS. With several k parts the rows meet by an atomic maximum over a filled Erow.
__global__ void __launch_bounds__(1024) k_row_exp_cols(const i32* __restrict__ X, const i64* __restrict__ e, int L,
                                                       i64 plane, i64 nrows, i64 rows, i64 K, i64 ldb, i64 ldr,
                                                       i64 ldk, i64 kchunk, i64* __restrict__ Erow) {
    __shared__ i64 part[16][64];
    const int lane = threadIdx.x & 63, kg = threadIdx.x >> 6;
    const i64 R = blockIdx.x * (i64)64 + lane;
    const i64 kb = (i64)blockIdx.y * kchunk, ke = kb + kchunk < K ? kb + kchunk : K;
    i64 mx = EZV;
    if (R < nrows) {
        const i64 b = R / rows, r = R % rows;
        const i64 base = b * ldb + r * ldr;
        const i32* top = X + (i64)(L - 1) * plane;
        for (i64 k0 = kb + kg; k0 < ke; k0 += 8 * 16) {
            i64 ex[8];
            u32 tw[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const i64 k = k0 + u * 16;
                ex[u] = k < ke ? e[base + k * ldk] : EZV;
                tw[u] = k < ke ? (u32)top[base + k * ldk] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const i64 t = entry_top_fast(X, plane, L, base + (k0 + u * 16) * ldk, ex[u], tw[u]);
                mx = t > mx ? t : mx;
            }
        }
    }
    part[kg][lane] = mx;
    __syncthreads();
    if (kg == 0 && R < nrows) {
        for (int q = 1; q < 16; q++) mx = part[q][lane] > mx ? part[q][lane] : mx;
        const i64 v = row_exp_of(mx, L);               // monotone in mx: the maximum commutes
        if (gridDim.y == 1) Erow[R] = v;
        else if (v != EZV) atomicMax((long long*)&Erow[R], (long long)v);
    }
}

// strided rows, few of them: 16 rows per block (lane = 16 rows x 4 k), its 16 waves over k, so that
// few rows still give enough blocks without splitting k (no fill, no atomics); the 64 k partial
// maxima of a row meet by lane exchange within the wave and in LDS across the waves
__global__ void __launch_bounds__(1024) k_row_exp_cols16(const i32* __restrict__ X, const i64* __restrict__ e, int L,
                                                         i64 plane, i64 nrows, i64 rows, i64 K, i64 ldb, i64 ldr,
                                                         i64 ldk, i64* __restrict__ Erow) {
    __shared__ i64 part[16][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rl = lane & 15, kq = lane >> 4;
    const i64 R = blockIdx.x * (i64)16 + rl;
    i64 mx = EZV;
    if (R < nrows) {
        const i64 b = R / rows, r = R % rows;
        const i64 base = b * ldb + r * ldr;
        const i32* top = X + (i64)(L - 1) * plane;
        for (i64 k0 = wave * 4 + kq; k0 < K; k0 += 8 * 64) {
            i64 ex[8];
            u32 tw[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const i64 k = k0 + u * 64;
                ex[u] = k < K ? e[base + k * ldk] : EZV;
                tw[u] = k < K ? (u32)top[base + k * ldk] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const i64 t = entry_top_fast(X, plane, L, base + (k0 + u * 64) * ldk, ex[u], tw[u]);
                mx = t > mx ? t : mx;
            }
        }
    }
    for (int o = 16; o <= 32; o <<= 1) {                  // the 4 k lanes of a row within the wave
        const i64 v = __shfl_xor(mx, o, 64);
        mx = v > mx ? v : mx;
    }
    if (kq == 0) part[wave][rl] = mx;
    __syncthreads();
    if (wave == 0 && kq == 0 && R < nrows) {
        for (int q = 1; q < 16; q++) mx = part[q][rl] > mx ? part[q][rl] : mx;
        Erow[R] = row_exp_of(mx, L);
    }
}

__global__ void __launch_bounds__(256) k_fill_i64(i64* __restrict__ p, i64 n, i64 v) {
    const i64 i = blockIdx.x * (i64)256 + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblocks(i64 n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace

extern "C" {

int sdpv_version(void) { return 1; }

#ifndef SDPV_SOURCE_HASH
#error "build with -DSDPV_SOURCE_HASH=<hash of the sources> (sdpvibe_amd/build.py does)"
#endif
uint64_t sdpv_source_hash(void) { return (uint64_t)SDPV_SOURCE_HASH; }

void* sdpv_host_alloc(int64_t bytes) {
    if (bytes <= 0) return nullptr;
    void* p = nullptr;
    if (hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

int sdpv_host_free(void* p) {
    if (!p) return SDPV_OK;
    return hipHostFree(p) == hipSuccess ? SDPV_OK : SDPV_EARG;
}

int sdpv_shift_round(const int32_t* X, int64_t Lin, int64_t B, const int64_t* s, int32_t* out, int64_t Lout,
                     int32_t* flag, int32_t per_entry, void* stream) {
    if (B < 0 || Lin < 1 || Lout < 1) return SDPV_EARG;
    if (B == 0) return SDPV_OK;
    if (!X || !s || !out || !flag) return SDPV_EARG;
    hipLaunchKernelGGL(k_shift_round, dim3(nblocks(B, TPB)), dim3(TPB), 0, as_stream(stream), X, (i64)B, (int)Lin,
                       s, out, (int)Lout, flag, (int)per_entry);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_add_same(const int32_t* A, const int32_t* Bm, int32_t* out, int64_t L, int64_t B, int32_t sgn,
                  void* stream) {
    if (B < 0 || L < 1 || (sgn != 1 && sgn != -1)) return SDPV_EARG;
    if (B == 0) return SDPV_OK;
    if (!A || !Bm || !out) return SDPV_EARG;
    hipLaunchKernelGGL(k_add_same, dim3(nblocks(B, TPB)), dim3(TPB), 0, as_stream(stream), A, Bm, out, (int)L,
                       (i64)B, (int)(sgn < 0));
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_bitlen(const int32_t* X, int64_t L, int64_t B, int64_t* out, void* stream) {
    if (B < 0 || L < 1) return SDPV_EARG;
    if (B == 0) return SDPV_OK;
    if (!X || !out) return SDPV_EARG;
    hipLaunchKernelGGL(k_bitlen, dim3(nblocks(B, TPB)), dim3(TPB), 0, as_stream(stream), X, (int)L, (i64)B, out);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mpf_top(const int32_t* m, const int64_t* e, int64_t L, int64_t n, int64_t* top, void* stream) {
    if (n < 0 || L < 1) return SDPV_EARG;
    if (n == 0) return SDPV_OK;
    if (!m || !e || !top) return SDPV_EARG;
    hipLaunchKernelGGL(k_mpf_top, dim3(nblocks(n, TPB)), dim3(TPB), 0, as_stream(stream), m, e, (int)L, (i64)n, top);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_badd_align(const int32_t* A, const int32_t* Bm, const int64_t* EA, const int64_t* EB, int32_t* out,
                    int64_t* EO, int64_t L, int64_t nb, int64_t rows, int64_t cols, int32_t sgn, void* stream) {
    if (L < 1 || nb < 0 || rows < 0 || cols < 0 || (sgn != 1 && sgn != -1)) return SDPV_EARG;
    const i64 entries = nb * rows * cols;
    if (entries == 0) return SDPV_OK;
    if (!A || !Bm || !EA || !EB || !out || !EO) return SDPV_EARG;
    hipLaunchKernelGGL(k_badd_align, dim3(nblocks(entries, TPB)), dim3(TPB), 0, as_stream(stream), A, Bm, EA, EB,
                       out, EO, (int)L, entries, (i64)rows, (i64)cols, (int)(sgn < 0));
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mpf_norm(const int32_t* X, const int64_t* eI, int64_t LI, int32_t* O, int64_t* eO, int64_t L, int64_t n,
                  void* stream) {
    if (n < 0 || L < 1 || LI < 1) return SDPV_EARG;
    if (n == 0) return SDPV_OK;
    if (!X || !eI || !O || !eO) return SDPV_EARG;
    hipLaunchKernelGGL(k_mpf_norm, dim3(nblocks(n, TPB)), dim3(TPB), 0, as_stream(stream), X, eI, (int)LI, O, eO,
                       (int)L, (i64)n);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mpf_mul(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB, int32_t* O,
                 int64_t* eO, int64_t L, int64_t n, int32_t bscalar, void* stream) {
    if (n < 0 || L < 1) return SDPV_EARG;
    if (L > 128) return SDPV_ESIZE;
    if (n == 0) return SDPV_OK;
    if (!A || !eA || !Bm || !eB || !O || !eO) return SDPV_EARG;
    const size_t lds = (size_t)4 * L * MUL_TPB * sizeof(u32);
    hipLaunchKernelGGL(k_mpf_mul, dim3(nblocks(n, MUL_TPB)), dim3(MUL_TPB), lds, as_stream(stream), A, eA, Bm, eB,
                       O, eO, (int)L, (i64)n, (int)(bscalar != 0));
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mpf_add(const int32_t* A, const int64_t* eA, const int32_t* Bm, const int64_t* eB, int32_t* O,
                 int64_t* eO, int64_t L, int64_t n, int32_t sgn, void* stream) {
    if (n < 0 || L < 1 || (sgn != 1 && sgn != -1)) return SDPV_EARG;
    if (L > 256) return SDPV_ESIZE;
    if (n == 0) return SDPV_OK;
    if (!A || !eA || !Bm || !eB || !O || !eO) return SDPV_EARG;
    const size_t lds = (size_t)(L + 2) * ADD_TPB * sizeof(u32);
    hipLaunchKernelGGL(k_mpf_add, dim3(nblocks(n, ADD_TPB)), dim3(ADD_TPB), lds, as_stream(stream), A, eA, Bm, eB,
                       O, eO, (int)L, (i64)n, (int)sgn);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_row_exp(const int32_t* X, const int64_t* e, int64_t L, int64_t plane, int64_t nb, int64_t rows, int64_t K,
                 int64_t ldb, int64_t ldr, int64_t ldk, int64_t* Erow, void* stream) {
    if (L < 1 || nb < 0 || rows < 0 || K < 0) return SDPV_EARG;
    const i64 nrows = nb * rows;
    if (nrows == 0) return SDPV_OK;
    if (!X || !e || !Erow) return SDPV_EARG;
    hipStream_t st = as_stream(stream);
    if (ldk == 1) {
        hipLaunchKernelGGL(k_row_exp_wave, dim3(nblocks(nrows, 4)), dim3(256), 0, st, X, e, (int)L, (i64)plane, nrows,
                           (i64)rows, (i64)K, (i64)ldb, (i64)ldr, Erow);
    } else {
        // few row groups of 64: 16-row blocks when they are enough (32 or more, or a short k),
        // else split k as well, so that the launch fills the device
        const i64 groups = (nrows + 63) / 64;
        if (groups < 128 && K > 256 && ((nrows + 15) / 16 >= 32 || K <= 4096)) {
            hipLaunchKernelGGL(k_row_exp_cols16, dim3(nblocks(nrows, 16)), dim3(1024), 0, st, X, e, (int)L,
                               (i64)plane, nrows, (i64)rows, (i64)K, (i64)ldb, (i64)ldr, (i64)ldk, Erow);
            SDPV_TRY_LAUNCH();
            return SDPV_OK;
        }
        i64 parts = 1;
        if (groups < 128 && K > 256) {
            parts = (256 + groups - 1) / groups;
            const i64 kmax = (K + 127) / 128;
            parts = parts < kmax ? parts : kmax;
            if (parts > 65535) parts = 65535;
        }
        const i64 kchunk = (K + parts - 1) / parts;
        if (parts > 1)
            hipLaunchKernelGGL(k_fill_i64, dim3(nblocks(nrows, 256)), dim3(256), 0, st, Erow, nrows, (i64)EZV);
        hipLaunchKernelGGL(k_row_exp_cols, dim3((unsigned)groups, (unsigned)parts), dim3(1024), 0, st, X, e, (int)L,
                           (i64)plane, nrows, (i64)rows, (i64)K, (i64)ldb, (i64)ldr, (i64)ldk, kchunk, Erow);
    }
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

}  // extern "C"
