// Exact products of a batch of fixed-point matrices with one or two vectors, formed directly on
// gfx950 without residues. A product with a narrow right-hand side (the forward and backward
// substitutions of a triangular solve with one column) costs little arithmetic but, through the
// residue path, a row-exponent pass, two splits, a product and a reconstruction; here it is one
// or two kernels whose output is the same exact integer (and bit length) that the reconstruction
// gives, so the rounding that follows, and every result, is unchanged.
//
// Entry (b, i, j): S = sum_k A[b, i, k] B[b, j, k] over L-word two's-complement operands. One wave
// takes one entry and one slice of k (slice_len); lane c (and c + 64) owns column c of
// the word products: it sums the unsigned products a_u b_v, u + v = c, as 32-bit halves, plus the
// corrections that turn the unsigned reading of the two top words into the signed one
// (a = a_unsigned - s_a 2^(32 L)). k runs in chunks of 64 staged in LDS (rows of 65 words, so that
// lanes reading one chunk column at different words use different banks), four k per step so
// that eight LDS reads are in flight. With one slice the wave carries its column sums into NW
// words itself; with several, each slice writes its column sums and k_exact_finish adds them.
#include "common.hpp"

namespace {

constexpr int DW = 4;        // waves per workgroup
constexpr int KC = 64;       // k per chunk
constexpr int KS = KC + 1;   // LDS row stride in words
constexpr int KSL = 256;     // longest slice of k

// k per slice for E entries and inner dimension K: slices of at most KSL values; shorter ones
// (multiples of the chunk, at least one chunk) while the launch would have fewer than about 1024
// waves. The sums are exact, so the partition changes no result.
__host__ inline i64 slice_len(i64 E, i64 K) {
    if (K <= KC) return K > 0 ? K : 1;
    i64 ns = (K + KSL - 1) / KSL;
    const i64 want = (1024 + E - 1) / E, nmax = (K + KC - 1) / KC;
    if (want > ns) ns = want < nmax ? want : nmax;
    const i64 len = (K + ns - 1) / ns;
    return (len + KC - 1) / KC * KC;
}

// column sums (lo, hi) of one entry over k0 <= k < k1, columns lane and lane + 64
__device__ __forceinline__ void column_sums(const i32* a, i64 planeA, i64 ldAk, const i32* bv, i64 planeB, i64 ldBk,
                                            int L, i64 k0, i64 k1, u32* As, u32* Bs, int lane, i64* lo, i64* hi) {
    for (i64 kb = k0; kb < k1; kb += KC) {
        const int kc = (int)(k1 - kb < KC ? k1 - kb : KC);
        wave_sync();                               // the previous chunk has been read
        for (int u = 0; u < L; u++) {
            As[u * KS + lane] = lane < kc ? (u32)a[u * planeA + (kb + lane) * ldAk] : 0u;
            Bs[u * KS + lane] = lane < kc ? (u32)bv[u * planeB + (kb + lane) * ldBk] : 0u;
        }
        wave_sync();
        const int kr = (kc + 3) & ~3;              // words beyond kc are zero
        for (int h = 0; h < 2; h++) {
            const int c = lane + 64 * h;
            if (c > 2 * L) continue;
            const int u0 = c - (L - 1) > 0 ? c - (L - 1) : 0, u1 = c < L - 1 ? c : L - 1;
            i64 sl = 0, sh = 0;
            for (int u = u0; u <= u1; u++) {
                const u32* ar = As + u * KS;
                const u32* br = Bs + (c - u) * KS;
                for (int kk = 0; kk < kr; kk += 4) {
                    u32 x[4], y[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        x[t] = ar[kk + t];
                        y[t] = br[kk + t];
                    }
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const u64 q = (u64)x[t] * y[t];
                        sl += (i64)(u32)q;
                        sh += (i64)(q >> 32);
                    }
                }
            }
            // a b = a_u b_u - s_a b_u 2^(32L) - s_b a_u 2^(32L) + s_a s_b 2^(64L)
            if (c >= L) {
                const u32* at = As + (L - 1) * KS;
                const u32* bt = Bs + (L - 1) * KS;
                const u32* aw = As + (c - L) * KS;
                const u32* bw = Bs + (c - L) * KS;
                for (int kk = 0; kk < kc; kk++) {
                    const u32 sa = at[kk] >> 31, sb = bt[kk] >> 31;
                    if (c < 2 * L) sl -= (i64)(sa ? bw[kk] : 0u) + (i64)(sb ? aw[kk] : 0u);
                    else sl += (i64)(sa & sb);
                }
            }
            lo[h] += sl;
            hi[h] += sh;
        }
    }
}

// the value sum_c (lo_c + hi_c 2^32) 2^(32 c) (columns 0..2L on lanes c, c - 64) as NW words
// W[w * E + e] and its bit length (of ~S when S < 0); every lane follows the carry chain
__device__ __forceinline__ void finish(const i64* lo, const i64* hi, int L, int NW, int lane, i64 e, i64 E,
                                       i32* __restrict__ W, i64* __restrict__ bl) {
    auto take = [&](const i64* v, int s) -> i64 {          // v of column s, 0 outside 0..2L
        if (s < 0 || s > 2 * L) return 0;
        const int src = s & 63;
        const i64 x0 = (i64)(((u64)(u32)__shfl((int)(u32)((u64)v[0] >> 32), src, 64) << 32) |
                             (u32)__shfl((int)(u32)v[0], src, 64));
        const i64 x1 = (i64)(((u64)(u32)__shfl((int)(u32)((u64)v[1] >> 32), src, 64) << 32) |
                             (u32)__shfl((int)(u32)v[1], src, 64));
        return s < 64 ? x0 : x1;
    };
    i64 carry = 0;
    u32 word[2] = {0u, 0u};
    for (int s = 0; s < NW; s++) {
        const i64 t = take(lo, s) + take(hi, s - 1) + carry;
        carry = t >> 32;
        if ((s & 63) == lane) word[s >> 6] = (u32)t;
    }
    const bool neg = (__shfl((int)word[(NW - 1) >> 6], (NW - 1) & 63, 64) >> 31) != 0;
    int top = -1;
    u32 wt = 0u;
    for (int h = 1; h >= 0 && top < 0; h--) {
        const int s = lane + 64 * h;
        const u32 x = s < NW ? (neg ? ~word[h] : word[h]) : 0u;
        const u64 nz = __ballot(x != 0u);
        if (nz) {
            const int l = 63 - __builtin_clzll(nz);
            top = l + 64 * h;
            wt = (u32)__shfl((int)x, l, 64);
        }
    }
    for (int h = 0; h < 2; h++) {
        const int s = lane + 64 * h;
        if (s < NW) W[(i64)s * E + e] = (i32)word[h];
    }
    if (lane == 0) bl[e] = top < 0 ? 0 : 32 * (i64)top + 32 - clz32(wt);
}

// one wave per (entry, slice); nslices == 1: words out directly, else column sums to Pl / Ph
// ((slice * E + e) * CW + c, CW = 2L + 1)
__global__ void __launch_bounds__(64 * DW) k_exact_dot(const i32* __restrict__ A, i64 planeA, i64 ldAb, i64 ldAr,
                                                      i64 ldAk, const i32* __restrict__ B, i64 planeB, i64 ldBb,
                                                      i64 ldBr, i64 ldBk, int L, i64 M, i64 N, i64 K, i64 E,
                                                      int nslices, i64 ksl, i32* __restrict__ W, int NW,
                                                      i64* __restrict__ bl, i64* __restrict__ Pl,
                                                      i64* __restrict__ Ph) {
    SDPV_DYNAMIC_LDS(u32, lds);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const i64 task = (i64)blockIdx.x * DW + wave;
    if (task >= E * nslices) return;               // uniform per wave; no workgroup barrier below
    const i64 e = task % E, sl = task / E;
    const i64 b = e / (M * N), rem = e - b * (M * N), i = rem / N, j = rem - i * N;
    u32* As = lds + (i64)wave * 2 * L * KS;
    u32* Bs = As + (i64)L * KS;
    i64 lo[2] = {0, 0}, hi[2] = {0, 0};
    const i64 k0 = sl * ksl, k1 = (k0 + ksl < K) ? k0 + ksl : K;
    column_sums(A + b * ldAb + i * ldAr, planeA, ldAk, B + b * ldBb + j * ldBr, planeB, ldBk, L,
                nslices == 1 ? 0 : k0, nslices == 1 ? K : k1, As, Bs, lane, lo, hi);
    if (nslices == 1) {
        finish(lo, hi, L, NW, lane, e, E, W, bl);
        return;
    }
    const int CW = 2 * L + 1;
    for (int h = 0; h < 2; h++) {
        const int c = lane + 64 * h;
        if (c < CW) {
            Pl[(sl * E + e) * CW + c] = lo[h];
            Ph[(sl * E + e) * CW + c] = hi[h];
        }
    }
}

// the slices' column sums added (exactly: each stays far inside 64 bits), then the words
__global__ void __launch_bounds__(64 * DW) k_exact_finish(const i64* __restrict__ Pl, const i64* __restrict__ Ph,
                                                         int L, i64 E, int nslices, i32* __restrict__ W, int NW,
                                                         i64* __restrict__ bl) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const i64 e = (i64)blockIdx.x * DW + wave;
    if (e >= E) return;
    const int CW = 2 * L + 1;
    i64 lo[2] = {0, 0}, hi[2] = {0, 0};
    for (int h = 0; h < 2; h++) {
        const int c = lane + 64 * h;
        if (c >= CW) continue;
        for (int s = 0; s < nslices; s++) {
            lo[h] += Pl[((i64)s * E + e) * CW + c];
            hi[h] += Ph[((i64)s * E + e) * CW + c];
        }
    }
    finish(lo, hi, L, NW, lane, e, E, W, bl);
}

}  // namespace

extern "C" {

int64_t sdpv_exact_dot_work(int64_t L, int64_t nb, int64_t M, int64_t N, int64_t K) {
    if (L < 1 || nb < 0 || M < 0 || N < 0 || K < 0) return SDPV_EARG;
    const i64 E = nb * M * N, ns = (K + slice_len(E, K) - 1) / slice_len(E, K);
    return ns <= 1 ? 0 : 2 * ns * E * (2 * L + 1);
}

int sdpv_exact_dot(const int32_t* A, int64_t planeA, int64_t ldAb, int64_t ldAr, int64_t ldAk, const int32_t* B,
                   int64_t planeB, int64_t ldBb, int64_t ldBr, int64_t ldBk, int64_t L, int64_t nb, int64_t M,
                   int64_t N, int64_t K, int32_t* W, int64_t NW, int64_t* bitlen, int64_t* work, int64_t nwork,
                   void* stream) {
    if (L < 1 || L > 63 || nb < 0 || M < 0 || N < 0 || K < 0) return SDPV_EARG;
    if (NW < 2 * L + 2 || NW > 128) return SDPV_EARG;
    const i64 E = nb * M * N;
    if (E == 0) return SDPV_OK;
    if (!W || !bitlen || (K > 0 && (!A || !B))) return SDPV_EARG;
    const i64 ksl = slice_len(E, K), ns = K > 0 ? (K + ksl - 1) / ksl : 1;
    const i64 need = ns > 1 ? 2 * ns * E * (2 * L + 1) : 0;
    if (need > 0 && (!work || nwork < need)) return SDPV_EARG;
    const i64 grid = (E * ns + DW - 1) / DW;
    if (grid > 0x7fffffffLL) return SDPV_ESIZE;
    const size_t lds = (size_t)DW * 2 * L * KS * sizeof(u32);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_exact_dot, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    hipStream_t st = as_stream(stream);
    i64* Pl = need ? work : nullptr;
    i64* Ph = need ? work + need / 2 : nullptr;
    hipLaunchKernelGGL(k_exact_dot, dim3((unsigned)grid), dim3(64 * DW), lds, st, A, (i64)planeA, (i64)ldAb,
                       (i64)ldAr, (i64)ldAk, B, (i64)planeB, (i64)ldBb, (i64)ldBr, (i64)ldBk, (int)L, (i64)M, (i64)N,
                       (i64)K, E, (int)ns, ksl, W, (int)NW, bitlen, Pl, Ph);
    SDPV_TRY_LAUNCH();
    if (ns > 1) {
        hipLaunchKernelGGL(k_exact_finish, dim3((unsigned)((E + DW - 1) / DW)), dim3(64 * DW), 0, st, Pl, Ph, (int)L,
                           E, (int)ns, W, (int)NW, bitlen);
        SDPV_TRY_LAUNCH();
    }
    return SDPV_OK;
}

}  // extern "C"
