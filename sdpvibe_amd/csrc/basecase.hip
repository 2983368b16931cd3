// Tile base case on gfx950: batched multiprecision Cholesky and triangular inverse of BT x BT
// tiles in wide fixed point (WL = L + 24 words, F = 32 L fractional bits), the arithmetic of the
// reference's chol_tiles / trinv_tiles (sdpvibe/basecase.py:13-246).
//
// Every product is the reference's truncated product mulfx (columns below LF - 2 dropped, signs
// applied to magnitudes), so results are bit-identical. Two forms of the product: one thread
// streams it word by word from LDS operands into its consumer (store, subtract, accumulate), or
// one wave holds a word per lane, sums the columns on the lanes and normalises them with shuffles
// and ballot carry-lookahead (lane_mulfx). When the tile's lower triangle fits in LDS
// (k_tile_chol_lds) the factorisation runs there and the inverse is built column-parallel
// (k_tile_inverse); otherwise k_tile_factor works on the tile in HBM.
#include <cstdlib>

#include "common.hpp"

namespace {

constexpr int TT = 256;  // threads per workgroup of k_tile_factor and k_tile_trsv

struct Slot {  // thread-strided word array in LDS
    u32* p;
    __device__ __forceinline__ u32& operator[](int k) const { return p[k * TT]; }
};

// copy |v| into slot s (WL words), return the sign of v
template <class Src>
__device__ __forceinline__ bool load_mag(Slot s, const Src& v, int WL) {
    const bool neg = (v(WL - 1) >> 31) != 0;
    u64 c = 1;
    for (int k = 0; k < WL; k++) {
        u32 w = v(k);
        if (neg) {
            const u64 t = (u64)(~w) + c;
            w = (u32)t;
            c = t >> 32;
        }
        s[k] = w;
    }
    return neg;
}

template <class S>
__device__ __forceinline__ int topl_slot(const S& s, int WL) {
    for (int i = WL - 1; i > 0; i--)
        if (s[i]) return i;
    return 0;
}

// column sum sum_{i=i0}^{i1} x[i] y[k - i] as 32-bit halves (lo: low words, hi: high words), eight
// terms at a time so that sixteen LDS reads are in flight before the multiplies use them
template <class X, class Y>
__device__ __forceinline__ void col_sum(const X& x, const Y& y, int k, int i0, int i1, u64& lo, u64& hi) {
    int i = i0;
    for (; i + 7 <= i1; i += 8) {
        u32 a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a[u] = x[i + u];
            b[u] = y[k - i - u];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const u64 p = (u64)a[u] * b[u];
            lo += (u32)p;
            hi += p >> 32;
        }
    }
    for (; i <= i1; i++) {
        const u64 p = (u64)x[i] * y[k - i];
        lo += (u32)p;
        hi += p >> 32;
    }
}

// Streams the magnitude words j = 0..WL-1 of mulfx(a, b, t) (sdpvibe/basecase.py:31-53) from the
// magnitudes x = |a|, y = |b| (word accessors x[i], y[i]); out(j, word). Columns kmin..LF-1 only
// feed the carry chain.
template <class X, class Y, class Out>
__device__ __forceinline__ void mulfx_mag(const X& x, const Y& y, int t, int WL, int LF, Out out) {
    const int ta = topl_slot(x, WL), tb = topl_slot(y, WL);
    const int la = t ? max(0, ta - t + 1) : 0, lb = t ? max(0, tb - t + 1) : 0;
    const int kmin = max(LF - 2, la + lb), kmax = ta + tb;
    u64 carry = 0, hprev = 0;
    auto column = [&](int k) -> u64 {
        u64 lo = 0, hi = 0;
        if (k <= kmax) col_sum(x, y, k, max(la, k - tb), min(ta, k - lb), lo, hi);
        const u64 sm = lo + hprev + carry;
        carry = sm >> 32;
        hprev = hi;
        return sm;
    };
    for (int k = kmin; k < LF && k <= kmax + 2; k++) column(k);
    for (int j = 0; j < WL; j++) {
        const int k = LF + j;
        out(j, (k >= kmin && k <= kmax + 2) ? (u32)column(k) : 0u);
    }
}

// the same product with the sign applied: out(j, word j of the two's complement result)
template <class Out>
__device__ __forceinline__ void mulfx_stream(Slot x, Slot y, bool neg, int t, int WL, int LF, Out out) {
    u64 nc = 1;
    mulfx_mag(x, y, t, WL, LF, [&](int j, u32 w) {
        if (neg) {
            const u64 tt = (u64)(~w) + nc;
            w = (u32)tt;
            nc = tt >> 32;
        }
        out(j, w);
    });
}

struct GlobalEntry {  // WL consecutive words in global memory
    const u32* p;
    __device__ __forceinline__ u32 operator()(int k) const { return p[k]; }
};
struct LdsEntry {
    const u32* p;
    __device__ __forceinline__ u32 operator()(int k) const { return p[k]; }
};

// ---------------------------------------------------------------- wave-level pivot arithmetic
// Arrays of WL words in LDS; called by all TT threads (threads >= 64 only meet the barriers).
struct Pivot {
    u32 *d, *r, *t1, *t2, *sx, *sy;
    u64* col;  // 2 * (2 WL + 4)
    int WL, LF;
};

__device__ __forceinline__ bool isneg_w(const u32* a, int WL) { return (a[WL - 1] >> 31) != 0; }

// r = mulfx(a, b, t) with the column sums spread over wave 0, the carry chain on lane 0
__device__ void wave_mul(const Pivot& P, u32* r, const u32* a, const u32* b, int t) {
    const int tid = threadIdx.x, WL = P.WL, LF = P.LF;
    __syncthreads();
    const bool sa = isneg_w(a, WL), sb = isneg_w(b, WL);
    if (tid == 0) {
        u64 ca = 1, cb = 1;
        for (int k = 0; k < WL; k++) {
            u32 x = a[k], y = b[k];
            if (sa) { const u64 q = (u64)(~x) + ca; x = (u32)q; ca = q >> 32; }
            if (sb) { const u64 q = (u64)(~y) + cb; y = (u32)q; cb = q >> 32; }
            P.sx[k] = x;
            P.sy[k] = y;
        }
    }
    __syncthreads();
    int ta = 0, tb = 0;
    for (int i = WL - 1; i > 0; i--)
        if (P.sx[i]) { ta = i; break; }
    for (int i = WL - 1; i > 0; i--)
        if (P.sy[i]) { tb = i; break; }
    const int la = t ? max(0, ta - t + 1) : 0, lb = t ? max(0, tb - t + 1) : 0;
    const int kmin = max(LF - 2, la + lb), kmax = ta + tb;
    if (tid < 64) {
        for (int k = kmin + tid; k <= kmax; k += 64) {
            u64 lo = 0, hi = 0;
            const int i0 = max(la, k - tb), i1 = min(ta, k - lb);
            for (int i = i0; i <= i1; i++) {
                const u64 p = (u64)P.sx[i] * P.sy[k - i];
                lo += (u32)p;
                hi += p >> 32;
            }
            P.col[2 * (k - kmin)] = lo;
            P.col[2 * (k - kmin) + 1] = hi;
        }
    }
    __syncthreads();
    if (tid == 0) {
        for (int j = 0; j < WL; j++) r[j] = 0;
        u64 carry = 0, hprev = 0;
        for (int k = kmin; k <= kmax + 2 && k < LF + WL; k++) {
            u64 lo = 0, hi = 0;
            if (k <= kmax) {
                lo = P.col[2 * (k - kmin)];
                hi = P.col[2 * (k - kmin) + 1];
            }
            const u64 sm = lo + hprev + carry;
            if (k >= LF) r[k - LF] = (u32)sm;
            carry = sm >> 32;
            hprev = hi;
        }
        if (sa ^ sb) {
            u64 c = 1;
            for (int j = 0; j < WL; j++) {
                const u64 q = (u64)(~r[j]) + c;
                r[j] = (u32)q;
                c = q >> 32;
            }
        }
    }
    __syncthreads();
}

__device__ double to_double_w(const u32* a, int WL, int LF) {  // basecase.py:59-68 (one thread)
    const bool s = isneg_w(a, WL);
    auto mag = [&](int k) {  // word k of |a|
        if (!s) return a[k];
        u64 c = 1;
        u32 w = 0;
        for (int q = 0; q <= k; q++) {
            const u64 t = (u64)(~a[q]) + c;
            w = (u32)t;
            c = t >> 32;
        }
        return w;
    };
    int t = 0;
    for (int k = WL - 1; k > 0; k--)
        if (mag(k)) { t = k; break; }
    double v = (double)mag(t);
    if (t >= 1) v = v * 4294967296.0 + (double)mag(t - 1);
    if (t >= 2) v = v * 4294967296.0 + (double)mag(t - 2);
    const int e = 32 * (t - 2 > 0 ? t - 2 : 0) - 32 * LF;
    v = ldexp(v, e);
    return s ? -v : v;
}

__device__ void from_double_w(u32* r, double v, int WL, int LF) {  // basecase.py:69-78
    for (int i = 0; i < WL; i++) r[i] = 0;
    if (!(v > 0.0) || !isfinite(v)) return;  // non-positive pivot: reciprocal 0
    int e;
    const double m = frexp(v, &e);
    u64 mi = (u64)ldexp(m, 53);
    int pos = e - 53 + 32 * LF;
    if (pos < 0) {
        mi = (-pos < 64) ? (mi >> (-pos)) : 0;
        pos = 0;
    }
    const int li = pos / 32, sh = pos % 32;
    u64 lo = mi << sh;
    u64 hi = sh ? (mi >> (64 - sh)) : 0;
    const u32 wds[4] = {(u32)lo, (u32)(lo >> 32), (u32)hi, (u32)(hi >> 32)};
    for (int k = 0; k < 4 && li + k < WL; k++) r[li + k] = wds[k];
}

// r = d^{-1/2}: double seed, Newton r <- r + r (1 - d r^2) / 2 with precision doubling
// (basecase.py:145-156)
__device__ void wave_rsqrt(const Pivot& P) {
    const int tid = threadIdx.x, WL = P.WL, LF = P.LF;
    if (tid == 0) from_double_w(P.r, 1.0 / sqrt(to_double_w(P.d, WL, LF)), WL, LF);
    for (int bits = 100;; bits *= 2) {
        int t = bits / 32 + 2;
        if (t >= WL) t = 0;
        wave_mul(P, P.t1, P.r, P.r, t);
        wave_mul(P, P.t2, P.t1, P.d, t);
        if (tid == 0) {  // t1 = one - t2
            i64 br = 0;
            for (int i = 0; i < WL; i++) {
                const i64 q = (i64)(i == LF ? 1u : 0u) - (i64)P.t2[i] + br;
                P.t1[i] = (u32)q;
                br = q >> 32;
            }
        }
        wave_mul(P, P.t2, P.r, P.t1, t);
        if (tid == 0) {  // r += t2 >> 1 (arithmetic)
            u64 c = 0;
            for (int i = 0; i < WL; i++) {
                const u32 hw = (i < WL - 1) ? ((P.t2[i] >> 1) | (P.t2[i + 1] << 31)) : (u32)(((int)P.t2[i]) >> 1);
                const u64 q = (u64)P.r[i] + hw + c;
                P.r[i] = (u32)q;
                c = q >> 32;
            }
        }
        __syncthreads();
        if (t == 0) break;
    }
}

// ---------------------------------------------------------------- lane-parallel pivot arithmetic
// For WL <= 64 the pivot chain runs on wave 0 with one word per lane (word k in lane k, lanes >= WL
// hold 0): magnitudes and additions use a carry lookahead over ballots, products sum their
// columns on the lanes and normalise the column sums with shuffles and one more lookahead. The
// results are word for word those of wave_mul / wave_rsqrt above (the same truncated product).
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ u32 lane_get(u32 v, int src) { return (u32)__shfl((int)v, src, 64); }
__device__ __forceinline__ u64 lane_get64(u64 v, int src) {
    return ((u64)lane_get((u32)(v >> 32), src) << 32) | lane_get((u32)v, src);
}
// value of the previous lane (0 in lane 0)
__device__ __forceinline__ u64 from_below64(u64 v) {
    const int l = lane_id();
    const u64 r = lane_get64(v, l > 0 ? l - 1 : 0);
    return l > 0 ? r : 0;
}

// top word index of the number (>= 1), else 0 (topl_slot semantics)
__device__ __forceinline__ int lane_top(u32 w, int WL) {
    const u64 nz = __ballot(lane_id() < WL && w != 0);
    return nz ? 63 - __builtin_clzll(nz) : 0;
}

// two's complement negation over WL words: ~w + 1, the +1 absorbed by the lowest nonzero word
__device__ u32 lane_neg(u32 w, int WL) {
    const int l = lane_id();
    const u64 nz = __ballot(l < WL && w != 0);
    const int first = nz ? __builtin_ctzll(nz) : 64;
    if (l >= WL) return 0u;
    return l < first ? 0u : (l == first ? ~w + 1u : ~w);
}

// a + b + cin mod 2^(32 WL): generate / propagate masks from two ballots, the carry into every word
// from one 64-bit addition ((G|P) + G + cin) ^ (G|P) ^ G
__device__ u32 lane_add(u32 a, u32 b, unsigned cin, int WL) {
    const int l = lane_id();
    const u64 s = (u64)a + b;
    const u64 G = __ballot(l < WL && (s >> 32) != 0);
    const u64 P = __ballot(l < WL && (u32)s == 0xffffffffu);
    const u64 X = G | P;
    const u64 carries = (X + G + cin) ^ X ^ G;
    return l < WL ? (u32)s + (u32)((carries >> l) & 1u) : 0u;
}

// mulfx(a, b, t) of two's complement lane numbers; xs, ys: WL words of LDS for the magnitudes
__device__ u32 lane_mulfx(u32 a, u32 b, int t, int WL, int LF, u32* xs, u32* ys) {
    const int l = lane_id();
    const bool sa = (lane_get(a, WL - 1) >> 31) != 0, sb = (lane_get(b, WL - 1) >> 31) != 0;
    const u32 ma = sa ? lane_neg(a, WL) : a, mb = sb ? lane_neg(b, WL) : b;
    const int ta = lane_top(ma, WL), tb = lane_top(mb, WL);
    __builtin_amdgcn_wave_barrier();                     // earlier readers of xs / ys are done
    if (l < WL) {
        xs[l] = ma;
        ys[l] = mb;
    }
    __builtin_amdgcn_wave_barrier();
    const int la = t ? max(0, ta - t + 1) : 0, lb = t ? max(0, tb - t + 1) : 0;
    const int kmin = max(LF - 2, la + lb), kmax = ta + tb;
    // column sums of columns kmin + l (h = 0) and kmin + 64 + l (h = 1) as lo / hi 32-bit halves
    u64 lo[2], hi[2];
    for (int h = 0; h < 2; h++) {
        const int k = kmin + l + 64 * h;
        u64 cl = 0, ch = 0;
        if (k <= kmax) col_sum(xs, ys, k, max(la, k - tb), min(ta, k - lb), cl, ch);
        lo[h] = cl;
        hi[h] = ch;
    }
    // V = sum_q (lo_q + hi_{q-1}) 2^(32 q), q = k - kmin over 128 positions (two per lane)
    // (every lane takes part in every shuffle; lane 0 of the upper half takes lane 63 of the lower)
    auto below = [&](u64 v0, u64 v1, u64& b0, u64& b1) {
        const u64 top = lane_get64(v0, 63), p0 = from_below64(v0), p1 = from_below64(v1);
        b0 = p0;
        b1 = l == 0 ? top : p1;
    };
    u64 b0, b1;
    below(hi[0], hi[1], b0, b1);
    const u64 s0 = lo[0] + b0, s1 = lo[1] + b1;
    // digits below 2^32 + 2^7: d_q = (s_q mod 2^32) + (s_{q-1} >> 32)
    below(s0 >> 32, s1 >> 32, b0, b1);
    const u64 d0 = (s0 & 0xffffffffull) + b0, d1 = (s1 & 0xffffffffull) + b1;
    // final digits: (d_q mod 2^32) + (d_{q-1} >> 32) + carry, the carries from a 128-bit lookahead
    below(d0 >> 32, d1 >> 32, b0, b1);
    const u64 e0 = (d0 & 0xffffffffull) + b0, e1 = (d1 & 0xffffffffull) + b1;
    const u64 G0 = __ballot((e0 >> 32) != 0), P0 = __ballot((u32)e0 == 0xffffffffu);
    const u64 G1 = __ballot((e1 >> 32) != 0), P1 = __ballot((u32)e1 == 0xffffffffu);
    const u64 X0 = G0 | P0, X1 = G1 | P1;
    const u64 sum0 = X0 + G0;
    const u64 C0 = sum0 ^ X0 ^ G0;
    const u64 C1 = (X1 + G1 + (sum0 < X0 ? 1u : 0u)) ^ X1 ^ G1;
    const u32 dig0 = (u32)e0 + (u32)((C0 >> l) & 1u), dig1 = (u32)e1 + (u32)((C1 >> l) & 1u);
    // output word j = digit of column LF + j
    const int q = LF + l - kmin;
    const u32 w0 = lane_get(dig0, q & 63), w1 = lane_get(dig1, q & 63);
    u32 w = (l >= WL || q < 0 || q > kmax + 2 - kmin) ? 0u : (q < 64 ? w0 : w1);
    __builtin_amdgcn_wave_barrier();
    return (sa ^ sb) ? lane_neg(w, WL) : w;
}

// basecase.py:59-68 on a lane number, the same value in every lane
__device__ double lane_to_double(u32 a, int WL, int LF) {
    const bool s = (lane_get(a, WL - 1) >> 31) != 0;
    const u32 m = s ? lane_neg(a, WL) : a;
    const int t = lane_top(m, WL);
    const u32 m0 = lane_get(m, t), m1 = lane_get(m, t >= 1 ? t - 1 : 0), m2 = lane_get(m, t >= 2 ? t - 2 : 0);
    double v = (double)m0;
    if (t >= 1) v = v * 4294967296.0 + (double)m1;
    if (t >= 2) v = v * 4294967296.0 + (double)m2;
    v = ldexp(v, 32 * (t - 2 > 0 ? t - 2 : 0) - 32 * LF);
    return s ? -v : v;
}

// basecase.py:69-78: the lane's word of the fixed-point value of v (0 for non-positive v)
__device__ u32 lane_from_double(double v, int WL, int LF) {
    const int l = lane_id();
    if (!(v > 0.0) || !isfinite(v) || l >= WL) return 0u;
    int e;
    const double m = frexp(v, &e);
    u64 mi = (u64)ldexp(m, 53);
    int pos = e - 53 + 32 * LF;
    if (pos < 0) {
        mi = (-pos < 64) ? (mi >> (-pos)) : 0;
        pos = 0;
    }
    const int li = pos / 32, sh = pos % 32;
    const u64 lo = mi << sh;
    const u64 hi = sh ? (mi >> (64 - sh)) : 0;
    const int k = l - li;
    return k == 0 ? (u32)lo : k == 1 ? (u32)(lo >> 32) : k == 2 ? (u32)hi : k == 3 ? (u32)(hi >> 32) : 0u;
}

// d^-1/2 (basecase.py:145-156): double seed, Newton r <- r + r (1 - d r^2) / 2, precision doubling
__device__ u32 lane_rsqrt(u32 d, int WL, int LF, u32* xs, u32* ys) {
    const int l = lane_id();
    u32 r = lane_from_double(1.0 / sqrt(lane_to_double(d, WL, LF)), WL, LF);
    for (int bits = 100;; bits *= 2) {
        int t = bits / 32 + 2;
        if (t >= WL) t = 0;
        u32 t1 = lane_mulfx(r, r, t, WL, LF, xs, ys);
        u32 t2 = lane_mulfx(t1, d, t, WL, LF, xs, ys);
        t1 = lane_add(l == LF ? 1u : 0u, ~t2, 1u, WL);            // one - t2
        t2 = lane_mulfx(r, t1, t, WL, LF, xs, ys);
        const u32 above = lane_get(t2, l < 63 ? l + 1 : 63);      // word l + 1
        const u32 half = (l < WL - 1) ? ((t2 >> 1) | (above << 31)) : (u32)(((int)t2) >> 1);
        r = lane_add(r, half, 0u, WL);                            // r += t2 >> 1 (arithmetic)
        if (t == 0) break;
    }
    return r;
}

// ---------------------------------------------------------------- the tile kernel
// W, X, S: AoS (nb, BT, BT, WL) words in global memory (W: working copy, factored in place;
// X: L^-1; S: products of the inverse). LDS: 2 operand slots per thread + pivot area + RINV.
__global__ void __launch_bounds__(TT) k_tile_factor(u32* __restrict__ Wg, u32* __restrict__ Xg, u32* __restrict__ Sg,
                                                   u32* __restrict__ Rg, int BT, int WL, int LF) {
    SDPV_DYNAMIC_LDS(u32, lds);
    const int tid = threadIdx.x;
    const i64 tile = blockIdx.x;
    u32* W = Wg + tile * BT * BT * WL;
    u32* X = Xg + tile * BT * BT * WL;
    u32* S = Sg + tile * BT * BT * WL;
    Slot sx{lds + tid}, sy{lds + WL * TT + tid};
    u32* piv = lds + 2 * WL * TT;
    Pivot P;
    P.d = piv;
    P.r = piv + WL;
    P.t1 = piv + 2 * WL;
    P.t2 = piv + 3 * WL;
    P.sx = piv + 4 * WL;
    P.sy = piv + 5 * WL;
    u32* rinv = piv + 6 * WL;                            // BT x WL
    P.col = (u64*)(rinv + BT * WL + ((BT * WL) & 1));    // 2 (2 WL + 4) u64 (8-byte aligned)
    P.WL = WL;
    P.LF = LF;
    auto ent = [&](u32* base, int i, int j) { return base + ((i64)i * BT + j) * WL; };

    for (int k = 0; k < BT; k++) {
        // pivot: d = A_kk, r = d^-1/2, A_kk = d r, RINV_k = r
        if (WL <= 64) {                                   // one word per lane on wave 0
            if (tid < 64) {
                const u32 d = tid < WL ? ent(W, k, k)[tid] : 0u;
                const u32 r = lane_rsqrt(d, WL, LF, P.sx, P.sy);
                const u32 dr = lane_mulfx(d, r, 0, WL, LF, P.sx, P.sy);
                if (tid < WL) {
                    ent(W, k, k)[tid] = dr;
                    rinv[k * WL + tid] = r;
                    Rg[(tile * BT + k) * WL + tid] = r;
                }
            }
            __syncthreads();
        } else {
            for (int q = tid; q < WL; q += TT) P.d[q] = ent(W, k, k)[q];
            __syncthreads();
            wave_rsqrt(P);
            wave_mul(P, P.t1, P.d, P.r, 0);
            for (int q = tid; q < WL; q += TT) {
                ent(W, k, k)[q] = P.t1[q];
                rinv[k * WL + q] = P.r[q];
                Rg[(tile * BT + k) * WL + q] = P.r[q];
            }
            __syncthreads();
        }
        // column scale: A_ik = mulfx(A_ik, r)
        for (int i = k + 1 + tid; i < BT; i += TT) {
            u32* aik = ent(W, i, k);
            const bool na = load_mag(sx, GlobalEntry{aik}, WL);
            load_mag(sy, LdsEntry{rinv + k * WL}, WL);
            mulfx_stream(sx, sy, na ^ isneg_w(rinv + k * WL, WL), 0, WL, LF, [&](int j, u32 w) { aik[j] = w; });
        }
        __syncthreads();
        // trailing update over lower pairs: A_ij -= mulfx(A_ik, A_jk)
        const int m = BT - k - 1, npairs = m * (m + 1) / 2;
        for (int p = tid; p < npairs; p += TT) {
            int ii = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
            while ((ii + 1) * (ii + 2) / 2 <= p) ii++;
            while (ii * (ii + 1) / 2 > p) ii--;
            const int jj = p - ii * (ii + 1) / 2;
            const int i = k + 1 + ii, j = k + 1 + jj;
            const bool na = load_mag(sx, GlobalEntry{ent(W, i, k)}, WL);
            const bool nb = load_mag(sy, GlobalEntry{ent(W, j, k)}, WL);
            u32* aij = ent(W, i, j);
            i64 br = 0;
            mulfx_stream(sx, sy, na ^ nb, 0, WL, LF, [&](int jw, u32 w) {
                const i64 q = (i64)aij[jw] - (i64)w + br;
                aij[jw] = (u32)q;
                br = q >> 32;
            });
        }
        __syncthreads();
    }
    // clear the upper triangle; X = 0
    for (int e = tid; e < BT * BT; e += TT) {
        const int i = e / BT, j = e % BT;
        for (int q = 0; q < WL; q++) {
            if (j > i) W[(i64)e * WL + q] = 0;
            X[(i64)e * WL + q] = 0;
        }
    }
    __syncthreads();
    // X = L^-1 row by row: x_ii = rinv_i, x_ik = -rinv_i * sum_{j=k}^{i-1} l_ij x_jk
    for (int i = 0; i < BT; i++) {
        const int np = i * (i + 1) / 2;  // pairs k <= j < i
        for (int p = tid; p < np; p += TT) {
            int jj = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
            while ((jj + 1) * (jj + 2) / 2 <= p) jj++;
            while (jj * (jj + 1) / 2 > p) jj--;
            const int kk = p - jj * (jj + 1) / 2, j = jj;
            const bool na = load_mag(sx, GlobalEntry{ent(W, i, j)}, WL);
            const bool nb = load_mag(sy, GlobalEntry{ent(X, j, kk)}, WL);
            u32* s = ent(S, kk, j);
            mulfx_stream(sx, sy, na ^ nb, 0, WL, LF, [&](int jw, u32 w) { s[jw] = w; });
        }
        __syncthreads();
        for (int kk = tid; kk <= i; kk += TT) {
            u32* xik = ent(X, i, kk);
            if (kk == i) {
                for (int q = 0; q < WL; q++) xik[q] = rinv[i * WL + q];
                continue;
            }
            // acc = sum_j S[kk][j] (exact, wrapping) into slot sx, then x_ik = -(acc * rinv_i)
            u64 c = 0;
            for (int q = 0; q < WL; q++) {
                u64 sm = c;
                for (int j = kk; j < i; j++) sm += ent(S, kk, j)[q];
                sy[q] = (u32)sm;  // raw sum words in sy (two's complement)
                c = sm >> 32;
            }
            // magnitude of acc into sx
            const bool na = (sy[WL - 1] >> 31) != 0;
            u64 cc = 1;
            for (int q = 0; q < WL; q++) {
                u32 w = sy[q];
                if (na) { const u64 tq = (u64)(~w) + cc; w = (u32)tq; cc = tq >> 32; }
                sx[q] = w;
            }
            load_mag(sy, LdsEntry{rinv + i * WL}, WL);
            const bool neg = !(na ^ isneg_w(rinv + i * WL, WL));  // negated product
            mulfx_stream(sx, sy, neg, 0, WL, LF, [&](int jw, u32 w) { xik[jw] = w; });
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- LDS-resident factorisation
// For tiles whose lower triangle fits in LDS (WL <= 64 at BT = 32): the same recurrences as
// k_tile_factor, with the working tile in LDS. Products are formed either by one thread (the
// column sums in a loop, operands read from LDS) or by one wave with a word per lane (lane_mulfx);
// both are the truncated product of the magnitudes with the sign applied, so the results are those
// of k_tile_factor bit for bit (A - (-|p|) is A + |p| modulo 2^(32 WL)). The wave form has far
// less latency per product; the thread form does 256 products at once. The inverse is built by
// k_tile_inverse, one workgroup per (tile, column).
struct Row {  // consecutive words in LDS
    u32* p;
    __device__ __forceinline__ u32& operator[](int k) const { return p[k]; }
};

constexpr int TWAVES = TT / 64;

// words of the LDS-resident factorisation for BT, WL and nw waves (entries padded to an odd
// stride WS)
__host__ __device__ inline i64 lds_factor_words(int BT, int WL, int nw) {
    const i64 WS = WL | 1;
    return (i64)BT * (BT + 1) / 2 * WS + (i64)BT * WS + WL + 2 * BT + 128 * nw;
}

// row i of the packed lower triangle starts at entry i (i + 1) / 2
__device__ __forceinline__ int tri_base(int i) { return i * (i + 1) / 2; }

// index p of the pairs (i, j), 0 <= j <= i, in row order -> (i, j)
__device__ __forceinline__ void tri_pair(int p, int& i, int& j) {
    int r = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= p) r++;
    while (r * (r + 1) / 2 > p) r--;
    i = r;
    j = p - r * (r + 1) / 2;
}

// A -= s * mulfx(x, y, 0) on LDS words (s = +1, or -1 when `add`): x, y magnitudes with top words
// ta, tb, A a WL-word two's complement target. The column sums are formed four columns at a time:
// x[i] is read once per block and the four operand words y[kb + c - i] sit in a window of registers
// that shifts by one word per i, so a block costs two LDS reads per i instead of eight; the target
// words of a block are read before its columns are summed. The column sums, the carry chain and the
// target words equal those of mulfx_mag and its consumer (terms outside the tops are zero).
__device__ __forceinline__ void trail_update(const u32* x, const u32* y, int ta, int tb, int WL, int LF, u32* A,
                                             bool add) {
    const int kmin = max(LF - 2, 0), kmax = ta + tb, kend = LF + WL;
    u64 carry = 0, hprev = 0, cy = 0;  // product chain; carry of the addition
    i64 br = 0;                          // borrow of the subtraction
    for (int kb = kmin; kb < kend; kb += 4) {
        u32 tw[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int j = kb + c - LF;
            tw[c] = (j >= 0 && j < WL) ? A[j] : 0u;
        }
        u64 lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
        if (kb <= kmax) {
            const int I0 = max(0, kb - tb), I1 = min(ta, kb + 3);
            u32 w[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int j = kb + c - I0;
                w[c] = j <= tb ? y[j] : 0u;
            }
            for (int i = I0; i <= I1; i += 4) {
                u32 xa[4], yn[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int ii = i + u, jn = kb - ii - 1;
                    xa[u] = ii <= I1 ? x[ii] : 0u;
                    yn[u] = (jn >= 0 && jn <= tb) ? y[jn] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const u64 p = (u64)xa[u] * w[c];
                        lo[c] += (u32)p;
                        hi[c] += p >> 32;
                    }
                    w[3] = w[2];
                    w[2] = w[1];
                    w[1] = w[0];
                    w[0] = yn[u];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int k = kb + c;
            if (k >= kend) break;
            u32 wv = 0u;
            if (k <= kmax + 2) {
                const u64 sm = lo[c] + hprev + carry;
                carry = sm >> 32;
                hprev = hi[c];
                wv = (u32)sm;
            }
            if (k >= LF) {
                if (add) {
                    const u64 q = (u64)tw[c] + wv + cy;
                    A[k - LF] = (u32)q;
                    cy = q >> 32;
                } else {
                    const i64 q = (i64)tw[c] - (i64)wv + br;
                    A[k - LF] = (u32)q;
                    br = q >> 32;
                }
            }
        }
    }
}

// Pivot k: wave 0 computes r = d^-1/2 and d r (lane_rsqrt, lane_mulfx). Column scale: the waves
// take the entries i > k in turn, L_ik = lane_mulfx(A_ik, r), and keep |L_ik|, its top word and its
// sign for the trailing products. Trailing update over the pairs k < j <= i: wave 0 updates the
// next diagonal entry and goes on to pivot k + 1 (look-ahead), the other waves update the
// remaining pairs, one pair per thread while there are more than `lane_pairs`, else one per wave.
// NT threads: 256 (one wave per SIMD), 512 or 1024 (two or four waves per SIMD: more instruction
// streams per SIMD, and 448 or 960 threads for the trailing products).
template <int NT>
__global__ void __launch_bounds__(NT) k_tile_chol_lds(u32* __restrict__ Wg, u32* __restrict__ Rg, int BT, int WL,
                                                     int LF, int lane_pairs) {
    constexpr int NW = NT / 64;
    SDPV_DYNAMIC_LDS(u32, lds);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const i64 tile = blockIdx.x;
    const int WS = WL | 1;
    const int ne = BT * (BT + 1) / 2;
    u32* T = lds;                  // ne x WS: the lower triangle, row by row
    u32* CL = T + ne * WS;         // BT x WS: |L_ik| of the pivot column
    u32* RK = CL + BT * WS;        // WL: r of the pivot
    u32* SG = RK + WL;             // BT: sign of L_ik
    u32* TP = SG + BT;             // BT: top word of |L_ik|
    u32* PX = TP + BT + 128 * wave;  // 64 + 64 per wave: lane_mulfx operands
    u32* PY = PX + 64;
    u32* W = Wg + tile * BT * BT * WL;
    auto te = [&](int i, int j) { return T + (i64)(tri_base(i) + j) * WS; };
    const int total = BT * BT * WL;
    // the lower triangle into LDS, 16 independent loads per thread in flight
    for (int base = tid; base < total; base += 16 * NT) {
        u32 v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int idx = base + u * NT;
            v[u] = idx < total ? W[idx] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int idx = base + u * NT;
            if (idx < total) {
                const int e = idx / WL, q = idx - e * WL, i = e / BT, j = e - i * BT;
                if (j <= i) te(i, j)[q] = v[u];
            }
        }
    }
    __syncthreads();
    // pivot k on wave 0, from the (updated) diagonal entry d held one word per lane:
    // r = d^-1/2, A_kk = mulfx(d, r), RINV_k = r
    auto pivot = [&](int k, u32 d) {
        const u32 r = lane_rsqrt(d, WL, LF, PX, PY);
        const u32 dr = lane_mulfx(d, r, 0, WL, LF, PX, PY);
        if (lane < WL) {
            te(k, k)[lane] = dr;
            RK[lane] = r;
            Rg[(tile * BT + k) * WL + lane] = r;
        }
    };
    if (wave == 0) pivot(0, lane < WL ? te(0, 0)[lane] : 0u);
    __syncthreads();
    for (int k = 0; k < BT; k++) {
        // column scale, entry i on wave (i - k - 1) mod NW: L_ik = mulfx(A_ik, r)
        {
            const u32 r = lane < WL ? RK[lane] : 0u;
            for (int i = k + 1 + wave; i < BT; i += NW) {
                u32* a = te(i, k);
                const u32 lv = lane_mulfx(lane < WL ? a[lane] : 0u, r, 0, WL, LF, PX, PY);
                const bool neg = (lane_get(lv, WL - 1) >> 31) != 0;
                const u32 mag = neg ? lane_neg(lv, WL) : lv;
                const int top = lane_top(mag, WL);
                if (lane < WL) {
                    a[lane] = lv;
                    CL[(i64)i * WS + lane] = mag;
                }
                if (lane == 0) {
                    SG[i] = neg;
                    TP[i] = top;
                }
            }
        }
        __syncthreads();
        if (k == BT - 1) break;
        // trailing update over the lower pairs k < j <= i: A_ij -= L_ik L_jk. Pair 0, the next
        // diagonal entry, on wave 0, which then computes pivot k + 1 while the other waves update the
        // other pairs (one per wave while there are at most `lane_pairs`, else one per thread)
        const int m = BT - k - 1, npairs = m * (m + 1) / 2;
        if (wave == 0) {
            const int r = k + 1;
            const u32 x = lane < WL ? te(r, k)[lane] : 0u;
            const u32 prod = lane_mulfx(x, x, 0, WL, LF, PX, PY);
            const u32 d = lane_add(lane < WL ? te(r, r)[lane] : 0u, ~prod, 1u, WL);
            if (lane < WL) te(r, r)[lane] = d;
            pivot(r, d);
        } else if (npairs - 1 <= lane_pairs) {
            for (int p = wave; p < npairs; p += NW - 1) {
                int ii, jj;
                tri_pair(p, ii, jj);
                const int r = k + 1 + ii, c = k + 1 + jj;
                const u32 x = lane < WL ? te(r, k)[lane] : 0u, y = lane < WL ? te(c, k)[lane] : 0u;
                const u32 prod = lane_mulfx(x, y, 0, WL, LF, PX, PY);
                u32* aij = te(r, c);
                const u32 nv = lane_add(lane < WL ? aij[lane] : 0u, ~prod, 1u, WL);
                if (lane < WL) aij[lane] = nv;
            }
        } else {
            for (int p = 1 + tid - 64; p < npairs; p += NT - 64) {
                int ii, jj;
                tri_pair(p, ii, jj);
                const int r = k + 1 + ii, c = k + 1 + jj;
                trail_update(CL + (i64)r * WS, CL + (i64)c * WS, TP[r], TP[c], WL, LF, te(r, c), SG[r] ^ SG[c]);
            }
        }
        __syncthreads();
    }
    // the factor back, zeros above the diagonal
    for (int idx = tid; idx < total; idx += NT) {
        const int e = idx / WL, q = idx - e * WL, i = e / BT, j = e - i * BT;
        W[idx] = j <= i ? te(i, j)[q] : 0u;
    }
}

// words of k_tile_inverse's LDS for BT, WL and nw waves; the row prefetch takes BT WL <= 8 threads
__host__ __device__ inline i64 lds_inverse_words(int BT, int WL, int nw) {
    return 4 * (i64)BT * WL + (i64)nw * WL + 128 * nw;
}
constexpr int INV_PRE = 8;  // words of the next row each thread prefetches

// Column k of X = L^-1 for one tile (the recurrences of k_tile_factor's inverse):
// x_kk = rinv_k; for i > k, x_ik = -mulfx(sum_{j=k}^{i-1} mulfx(l_ij, x_jk), rinv_i), every product
// truncated, the sum exact (modulo 2^(32 WL)). Step i: the NT / 64 waves take the products j = k..i-1
// in turn, one lane_mulfx each, and add them into a running lane sum; wave 0 adds the waves' sums
// and finishes x_ik with one lane_mulfx. Row i + 1 of L is loaded while step i computes.
// Grid (nb, BT), NT threads. X entries above the diagonal are written as zeros.
template <int NT>
__global__ void __launch_bounds__(NT) k_tile_inverse(const u32* __restrict__ Wg, const u32* __restrict__ Rg,
                                                    u32* __restrict__ Xg, int BT, int WL, int LF) {
    constexpr int NW = NT / 64;
    SDPV_DYNAMIC_LDS(u32, lds);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const i64 tile = blockIdx.x;
    const int k = blockIdx.y;
    u32* XT = lds;                      // BT x WL: x_jk
    u32* RI = XT + BT * WL;             // BT x WL: rinv
    u32* LR = RI + BT * WL;             // 2 x BT x WL: rows of L (slot j holds l_ij), double buffered
    u32* PART = LR + 2 * BT * WL;       // NW x WL: the waves' partial sums
    u32* PX = PART + NW * WL + 128 * wave;
    u32* PY = PX + 64;
    const u32* W = Wg + tile * BT * BT * WL;
    const u32* R = Rg + tile * BT * WL;
    u32* X = Xg + tile * BT * BT * WL;
    for (int idx = tid; idx < BT * WL; idx += NT) RI[idx] = R[idx];
    for (int i = 0; i < k; i++)
        for (int q = tid; q < WL; q += NT) X[((i64)i * BT + k) * WL + q] = 0u;
    // row k + 1 into buffer 0 (entries j = k..k, at slot j)
    if (k + 1 < BT)
        for (int idx = tid; idx < WL; idx += NT) LR[(i64)k * WL + idx] = W[((i64)(k + 1) * BT + k) * WL + idx];
    __syncthreads();
    if (wave == 0 && lane < WL) {       // x_kk = r_k
        const u32 r = RI[k * WL + lane];
        X[((i64)k * BT + k) * WL + lane] = r;
        XT[k * WL + lane] = r;
    }
    __syncthreads();
    for (int i = k + 1; i < BT; i++) {
        u32* cur = LR + ((i - k - 1) & 1) * BT * WL;
        u32* nxt = LR + ((i - k) & 1) * BT * WL;
        // prefetch row i + 1: (i + 1 - k) WL words from entry (i + 1, k)
        u32 pre[INV_PRE];
        const int nw = i + 1 < BT ? (i + 1 - k) * WL : 0;
        const u32* src = W + ((i64)(i + 1) * BT + k) * WL;
#pragma unroll
        for (int u = 0; u < INV_PRE; u++) {
            const int idx = tid + u * NT;
            pre[u] = idx < nw ? src[idx] : 0u;
        }
        // this wave's products, summed exactly
        u32 acc = 0u;
        for (int j = k + wave; j < i; j += NW) {
            const u32 lij = lane < WL ? cur[j * WL + lane] : 0u;
            const u32 xjk = lane < WL ? XT[j * WL + lane] : 0u;
            acc = lane_add(acc, lane_mulfx(lij, xjk, 0, WL, LF, PX, PY), 0u, WL);
        }
        if (lane < WL) PART[wave * WL + lane] = acc;
#pragma unroll
        for (int u = 0; u < INV_PRE; u++) {
            const int idx = tid + u * NT;
            if (idx < nw) nxt[k * WL + idx] = pre[u];
        }
        __syncthreads();
        if (wave == 0) {
            u32 sum = lane < WL ? PART[lane] : 0u;
            for (int w = 1; w < NW; w++) sum = lane_add(sum, lane < WL ? PART[w * WL + lane] : 0u, 0u, WL);
            const u32 x = lane_neg(lane_mulfx(sum, lane < WL ? RI[i * WL + lane] : 0u, 0, WL, LF, PX, PY), WL);
            if (lane < WL) {
                X[((i64)i * BT + k) * WL + lane] = x;
                XT[i * WL + lane] = x;
            }
        }
        __syncthreads();
    }
}

// Forward substitution L y = v for many vectors with one factored tile (the reference's trsv_warp,
// sdpvibe/basecase.py:228-245): 32 lanes per vector, lane i owns v_i (BT <= 32), kept in place in
// V (nvec, BT, WL). Step j: lane j scales y_j = mulfx(v_j, rinv_j); after the barrier every lane
// i > j subtracts mulfx(l_ij, y_j). All threads run every step (barriers), idle lanes included.
__global__ void __launch_bounds__(TT) k_tile_trsv(const u32* __restrict__ Lw, const u32* __restrict__ R,
                                                 u32* __restrict__ V, i64 nvec, int BT, int WL, int LF) {
    SDPV_DYNAMIC_LDS(u32, lds);
    const int tid = threadIdx.x, lane = tid & 31;
    const i64 vec = (i64)blockIdx.x * (TT / 32) + (tid >> 5);
    const bool live = vec < nvec && lane < BT;
    Slot sx{lds + tid}, sy{lds + WL * TT + tid};
    u32* vb = V + vec * BT * WL;
    for (int j = 0; j < BT; j++) {
        if (live && lane == j) {
            u32* vj = vb + (i64)j * WL;
            const u32* rj = R + (i64)j * WL;
            const bool na = load_mag(sx, GlobalEntry{vj}, WL);
            const bool nr = load_mag(sy, GlobalEntry{rj}, WL);
            mulfx_stream(sx, sy, na ^ nr, 0, WL, LF, [&](int q, u32 w) { vj[q] = w; });
        }
        __syncthreads();
        if (live && lane > j) {
            u32* vi = vb + (i64)lane * WL;
            const bool na = load_mag(sx, GlobalEntry{Lw + ((i64)lane * BT + j) * WL}, WL);
            const bool ny = load_mag(sy, GlobalEntry{vb + (i64)j * WL}, WL);
            i64 br = 0;
            mulfx_stream(sx, sy, na ^ ny, 0, WL, LF, [&](int q, u32 w) {
                const i64 d = (i64)vi[q] - (i64)w + br;
                vi[q] = (u32)d;
                br = d >> 32;
            });
        }
        __syncthreads();
    }
}

// planes (L limbs, value at exponent efx) -> AoS wide words (value * 2^40 in WL words)
__global__ void __launch_bounds__(256) k_tile_widen(const i32* __restrict__ A, i64 plane, int L, int WL,
                                                    u32* __restrict__ W, i64 n) {
    const i64 e = blockIdx.x * (i64)256 + threadIdx.x;
    if (e >= n) return;
    struct Src {
        const i32* p;
        i64 s;
        __device__ __forceinline__ u32 operator()(int k) const { return (u32)p[(i64)k * s]; }
    } src{A + e, plane};
    u32* out = W + e * WL;
    shr_round(src, L, -40, WL, [&](int k, u32 w) { out[k] = w; });
}

// AoS wide -> planes: L (rounded, value / 2^40 into L limbs) and X (WL words)
__global__ void __launch_bounds__(256) k_tile_narrow(const u32* __restrict__ W, const u32* __restrict__ X, int L,
                                                     int WL, i32* __restrict__ Lout, i32* __restrict__ Xw,
                                                     i64 plane, i64 n, i32* __restrict__ flag) {
    const i64 e = blockIdx.x * (i64)256 + threadIdx.x;
    if (e >= n) return;
    const u32* w = W + e * WL;
    const int bad = shr_round([&](int k) { return w[k]; }, WL, 40, L,
                              [&](int k, u32 v) { Lout[(i64)k * plane + e] = (i32)v; });
    if (bad && flag) atomicOr(flag, 1);
    const u32* x = X + e * WL;
    for (int k = 0; k < WL; k++) Xw[(i64)k * plane + e] = (i32)x[k];
}

inline unsigned nblk(i64 n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// trailing updates with at most this many pairs run one pair per wave (SDPV_TILE_LANE_PAIRS)
inline int tile_lane_pairs() {
    static const int v = [] {
        const char* e = getenv("SDPV_TILE_LANE_PAIRS");
        return e ? atoi(e) : 32;
    }();
    return v;
}

// threads per (tile, column) of the inverse: 512 unless SDPV_INV_THREADS is 256 or 1024
inline int inverse_threads() {
    static const int v = [] {
        const char* e = getenv("SDPV_INV_THREADS");
        const int t = e ? atoi(e) : 512;
        return (t == 256 || t == 1024) ? t : 512;
    }();
    return v;
}

// threads of the LDS-resident factorisation: 1024 unless SDPV_TILE_THREADS is 256 or 512. Measured
// at L = 24: chol(Q) of N = 3000 in 0.241 s with 256 threads, 0.203 s with 512 and 0.197 s with 1024
// (all with the 512-thread inverse; 0.210 s with 512 and the 256-thread inverse).
inline int tile_threads() {
    static const int v = [] {
        const char* e = getenv("SDPV_TILE_THREADS");
        const int t = e ? atoi(e) : 1024;
        return (t == 256 || t == 512) ? t : 1024;
    }();
    return v;
}

}  // namespace

extern "C" {

int sdpv_tile_chol(const int32_t* A, int64_t L, int64_t nb, int64_t BT, int32_t* Lout, int32_t* Xw, uint32_t* work,
                   int32_t* flag, void* stream) {
    if (L < 1 || nb < 0 || BT < 1 || BT > 64) return SDPV_EARG;
    if (nb == 0) return SDPV_OK;
    if (!A || !Lout || !Xw || !work) return SDPV_EARG;
    const int WL = (int)L + 24;
    const size_t lds = (size_t)2 * WL * TT * 4 + (size_t)(6 * WL + BT * WL + 2) * 4 + (size_t)2 * (2 * WL + 4) * 8;
    if (lds > 160 * 1024) return SDPV_ESIZE;
    hipStream_t st = as_stream(stream);
    const i64 n = nb * BT * BT;
    u32* W = work;
    u32* X = work + n * WL;
    u32* S = work + 2 * n * WL;
    u32* RI = work + 3 * n * WL;
    hipLaunchKernelGGL(k_tile_widen, dim3(nblk(n, 256)), dim3(256), 0, st, A, n, (int)L, WL, W, n);
    SDPV_TRY_LAUNCH();
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_tile_factor, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_tile_chol_lds<256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_tile_chol_lds<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_tile_chol_lds<1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    const int nt = tile_threads();
    const size_t lds_f = (size_t)lds_factor_words((int)BT, WL, nt / 64) * 4;
    const int ni = inverse_threads();
    const size_t lds_i = (size_t)lds_inverse_words((int)BT, WL, ni / 64) * 4;
    if (WL <= 64 && lds_f <= 160 * 1024 && BT * WL <= INV_PRE * 256) {
        // the tile stays in LDS; the inverse one workgroup per column
        if (nt == 1024)
            hipLaunchKernelGGL(k_tile_chol_lds<1024>, dim3((unsigned)nb), dim3(1024), lds_f, st, W, RI, (int)BT, WL,
                               (int)L, tile_lane_pairs());
        else if (nt == 512)
            hipLaunchKernelGGL(k_tile_chol_lds<512>, dim3((unsigned)nb), dim3(512), lds_f, st, W, RI, (int)BT, WL,
                               (int)L, tile_lane_pairs());
        else
            hipLaunchKernelGGL(k_tile_chol_lds<256>, dim3((unsigned)nb), dim3(256), lds_f, st, W, RI, (int)BT, WL,
                               (int)L, tile_lane_pairs());
        SDPV_TRY_LAUNCH();
        if (ni == 1024)
            hipLaunchKernelGGL(k_tile_inverse<1024>, dim3((unsigned)nb, (unsigned)BT), dim3(1024), lds_i, st, W, RI,
                               X, (int)BT, WL, (int)L);
        else if (ni == 512)
            hipLaunchKernelGGL(k_tile_inverse<512>, dim3((unsigned)nb, (unsigned)BT), dim3(512), lds_i, st, W, RI, X,
                               (int)BT, WL, (int)L);
        else
            hipLaunchKernelGGL(k_tile_inverse<256>, dim3((unsigned)nb, (unsigned)BT), dim3(256), lds_i, st, W, RI, X,
                               (int)BT, WL, (int)L);
        SDPV_TRY_LAUNCH();
    } else {
        hipLaunchKernelGGL(k_tile_factor, dim3((unsigned)nb), dim3(TT), lds, st, W, X, S, RI, (int)BT, WL, (int)L);
        SDPV_TRY_LAUNCH();
    }
    hipLaunchKernelGGL(k_tile_narrow, dim3(nblk(n, 256)), dim3(256), 0, st, W, X, (int)L, WL, Lout, Xw, n, n, flag);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_tile_trsv(const uint32_t* Lw, const uint32_t* RINV, uint32_t* V, int64_t nvec, int64_t BT, int64_t WL,
                   int64_t LF, void* stream) {
    if (nvec < 0 || BT < 1 || BT > 32 || LF < 1 || WL < LF + 2) return SDPV_EARG;
    if (nvec == 0) return SDPV_OK;
    if (!Lw || !RINV || !V) return SDPV_EARG;
    const size_t lds = (size_t)2 * WL * TT * 4;
    if (lds > 160 * 1024) return SDPV_ESIZE;
    const i64 grid = (nvec + TT / 32 - 1) / (TT / 32);
    if (grid > 0x7fffffffLL) return SDPV_ESIZE;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_tile_trsv, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL(k_tile_trsv, dim3((unsigned)grid), dim3(TT), lds, as_stream(stream), Lw, RINV, V, nvec,
                       (int)BT, (int)WL, (int)LF);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

}  // extern "C"
