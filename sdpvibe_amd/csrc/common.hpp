// Shared device helpers for the gfx950 multiprecision kernels.
//
// Multiword integers are read through small accessor functors `src(i) -> u32 word i` so the same
// streaming routines serve limb planes in HBM, per-thread words in LDS and values being
// reconstructed. Every routine runs over a run-time word count: no per-L template instances.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/sdpvibe_hip.h"

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;
typedef int32_t i32;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

#define EZV (-(1LL << 40))

// Dynamic LDS of a kernel (the host emulator in tests/emu supplies its own definition).
#ifndef SDPV_DYNAMIC_LDS
#define SDPV_DYNAMIC_LDS(type, name) extern __shared__ __align__(16) type name[]
#endif
// 16 bytes per lane from global memory straight into LDS at ldsbase + 16 * lane (gfx950 direct
// load; ldsbase uniform across the wave), and a wait until at most n vector-memory operations of
// this wave are outstanding (s_waitcnt immediate: vmcnt split over bits 3:0 and 15:14).
#ifndef SDPV_LDS_DMA16
#define SDPV_LDS_DMA16(gptr, ldsbase)                                                            \
    __builtin_amdgcn_global_load_lds((const void*)(gptr),                                        \
                                     (__attribute__((address_space(3))) void*)(ldsbase), 16, 0, 0)
#define SDPV_WAIT_VMCNT(n) \
    __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | ((((n) >> 4) & 3) << 14))
#endif
#define SDPV_TRY_LAUNCH()                                   \
    do {                                                    \
        hipError_t err__ = hipGetLastError();               \
        if (err__ != hipSuccess) return SDPV_ELAUNCH;       \
    } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Orders LDS writes of a wave before later reads by other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchanges lanes 32-63 of x with lanes 0-31 of y (v_permlane32_swap). Writing a register as
// (lanes 0-31, lanes 32-63): (x_lo, x_hi), (y_lo, y_hi) become x = (x_lo, y_lo), y = (x_hi, y_hi).
__device__ __forceinline__ void swap_halves(u32& x, u32& y) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}

__device__ __forceinline__ int clz32(u32 x) { return __clz((int)x); }

// Word i of the infinite two's-complement extension of an ns-word value (sdpvibe/limbops.py:22-28).
template <class Src>
__device__ __forceinline__ u32 ext_word(const Src& src, int ns, u32 ext, i64 i) {
    if (i < 0) return 0u;
    if (i >= ns) return ext;
    return src((int)i);
}

// Streaming rounding shift, the semantics of limbops._shift_round_kernel (sdpvibe/limbops.py:32-62):
// out word k (k < Lout) of round_half_up(x / 2^s) = floor((x + 2^(s-1)) / 2^s) for s > 0, exact
// left shift for s <= 0, x two's complement of ns words. Returns 1 when word Lout of the rounded
// value differs from the sign extension of word Lout-1 (the reference's overflow test).
template <class Src, class Out>
__device__ __forceinline__ int shr_round(const Src& src, int ns, i64 s, int Lout, Out out) {
    const u32 ext = (src(ns - 1) >> 31) ? 0xffffffffu : 0u;
    const i64 q = s >> 5;
    const int r = (int)(s & 31);
    u64 carry = 0;
    if (s >= 1) {
        const i64 sb = s - 1;
        carry = (ext_word(src, ns, ext, sb >> 5) >> (int)(sb & 31)) & 1u;
    }
    u32 prev = 0;
    u64 lo = ext_word(src, ns, ext, q);
    for (int k = 0; k <= Lout; k++) {
        const u64 hi = ext_word(src, ns, ext, q + k + 1);   // each source word is read once
        const u64 v = (((lo | (hi << 32)) >> r) & 0xffffffffull) + carry;
        lo = hi;
        carry = v >> 32;
        const u32 w = (u32)v;
        if (k < Lout) {
            out(k, w);
            prev = w;
        } else {
            const u32 want = (prev >> 31) ? 0xffffffffu : 0u;
            return w != want;
        }
    }
    return 0;
}

// Words k < Lout of floor(v / 2^s) + c (s < 0: exact left shift), v the two's complement value
// of ns words read through src (zero below word 0, sign words above ns - 1), in batches of 8
// independent reads; returns word Lout of that sum. |s| is clamped to where the words no longer
// change (beyond the value, or beyond the output).
template <class Src, class Out>
__device__ __forceinline__ u32 shift_words(const Src& src, int ns, u32 ext, i64 s, u32 c, int Lout, Out out) {
    const i64 smax = 32 * (i64)ns + 32, smin = -32 * ((i64)Lout + 2);
    s = s > smax ? smax : (s < smin ? smin : s);
    auto wd = [&](int t) -> u32 {
        const u32 v = src(t < 0 ? 0 : (t >= ns ? ns - 1 : t));
        return t < 0 ? 0u : (t >= ns ? ext : v);
    };
    const int qs = (int)(s >> 5), r = (int)(s & 31);
    for (int k0 = 0; k0 < Lout; k0 += 8) {
        u32 x[9];
#pragma unroll
        for (int u = 0; u < 9; u++) x[u] = wd(qs + k0 + u);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (k0 + u < Lout) {
                const u32 d = (u32)((((u64)x[u + 1] << 32) | x[u]) >> r);
                const u32 t = d + c;
                c = t < d ? 1u : 0u;
                out(k0 + u, t);
            }
        }
    }
    return (u32)((((u64)wd(qs + Lout + 1) << 32) | wd(qs + Lout)) >> r) + c;
}

// shr_round with the source read in batches (shift_words): same words, same return value.
template <class Src, class Out>
__device__ __forceinline__ int shr_round_batched(const Src& src, int ns, i64 s, int Lout, Out out) {
    const u32 ext = (src(ns - 1) >> 31) ? 0xffffffffu : 0u;
    u32 c = 0;
    if (s >= 1) {
        const i64 sb = s - 1;
        c = (sb >> 5) >= ns ? (ext & 1u) : (src((int)(sb >> 5)) >> (int)(sb & 31)) & 1u;
    }
    u32 prev = 0;
    const u32 top = shift_words(src, ns, ext, s, c, Lout, [&](int k, u32 v) {
        out(k, v);
        prev = v;
    });
    return top != ((prev >> 31) ? 0xffffffffu : 0u);
}

// Bit length of v (of ~v when v < 0), limbops._bitlen_kernel (sdpvibe/limbops.py:197-210).
template <class Src>
__device__ __forceinline__ int bitlen_tc(const Src& src, int n) {
    const u32 ext = (src(n - 1) >> 31) ? 0xffffffffu : 0u;
    for (int k = n - 1; k >= 0; k--) {
        const u32 x = src(k) ^ ext;
        if (x) return 32 * k + 32 - clz32(x);
    }
    return 0;
}

// Magnitude view of an n-word two's-complement value: |x| word i, computed on the fly.
// For negative x with lowest nonzero word z: words below z are 0, word z is -x_z, above ~x_i.
template <class Src>
struct MagView {
    Src src;
    int n;
    bool neg;
    int z;
    __device__ __forceinline__ MagView(const Src& s, int n_) : src(s), n(n_) {
        neg = (src(n - 1) >> 31) != 0;
        z = 0;
        if (neg) {
            while (z < n - 1 && src(z) == 0u) z++;
        }
    }
    __device__ __forceinline__ u32 operator()(int i) const {
        if (!neg) return src(i);
        if (i < z) return 0u;
        const u32 v = src(i);
        return i == z ? (u32)(0u - v) : ~v;
    }
};

// Plain bit length of a magnitude (mpf.py:32-35 `bitlen`).
template <class Src>
__device__ __forceinline__ int bitlen_mag(const Src& src, int n) {
    for (int k = n - 1; k >= 0; k--) {
        const u32 x = src(k);
        if (x) return 32 * k + 32 - clz32(x);
    }
    return 0;
}

// shift_mag of sdpvibe/mpf.py:36-56: dst[0..nd) = round_half_up(|src| / 2^s) (s > 0, magnitude
// half up, carry out of the top word dropped) or |src| * 2^-s (s <= 0, bits above nd dropped).
template <class Src, class Out>
__device__ __forceinline__ void shift_mag(const Src& src, int ns, i64 s, int nd, Out out) {
    if (s > 0) {
        const i64 q = s >> 5;
        const int r = (int)(s & 31);
        const i64 sb = s - 1, qb = sb >> 5;
        const int rb = (int)(sb & 31);
        u64 c = (qb < ns) ? ((src((int)qb) >> rb) & 1u) : 0u;
        for (int k = 0; k < nd; k++) {
            const i64 i0 = q + k, i1 = q + k + 1;
            const u32 lo = i0 < ns ? src((int)i0) : 0u;
            const u32 hi = i1 < ns ? src((int)i1) : 0u;
            const u32 d = r ? ((lo >> r) | (hi << (32 - r))) : lo;
            const u64 t = (u64)d + c;
            c = t >> 32;
            out(k, (u32)t);
        }
    } else {
        const i64 t = -s, q = t >> 5;
        const int r = (int)(t & 31);
        for (int k = 0; k < nd; k++) {
            const i64 i0 = k - q, i1 = k - q - 1;
            const u32 lo = (i0 >= 0 && i0 < ns) ? src((int)i0) : 0u;
            const u32 lo2 = (i1 >= 0 && i1 < ns) ? src((int)i1) : 0u;
            out(k, r ? ((lo << r) | (lo2 >> (32 - r))) : lo);
        }
    }
}

// Streaming two's-complement negation of words produced in increasing order.
struct NegStream {
    bool neg;
    u64 c;
    __device__ __forceinline__ explicit NegStream(bool n) : neg(n), c(1) {}
    __device__ __forceinline__ u32 operator()(u32 v) {
        if (!neg) return v;
        const u64 t = (u64)(~v) + c;
        c = t >> 32;
        return (u32)t;
    }
};

// Exact residue of a signed 32-bit integer modulo m (< 2^15) in [0, m), via a double reciprocal
// and one correction each way (|x| < 2^31 keeps the quotient estimate within 1).
__device__ __forceinline__ int mod_m(int x, int m, double minv) {
    const int q = (int)floor((double)x * minv);
    int r = x - q * m;
    if (r < 0) r += m;
    if (r >= m) r -= m;
    return r;
}

// Centred residue in [-(m-1)/2, (m-1)/2] split into 128*hi + lo, lo in [-64, 63].
__device__ __forceinline__ void split_centred(int r, int m, int& hi, int& lo) {
    if (r > (m >> 1)) r -= m;
    lo = ((r + 64) & 127) - 64;
    hi = (r - lo) >> 7;
}

// Bytewise a + b of packed int8 lanes (no carries between bytes); exact when each byte sum
// stays in [-128, 127].
__device__ __forceinline__ u32 add_bytes(u32 a, u32 b) {
    return ((a & 0x7f7f7f7fu) + (b & 0x7f7f7f7fu)) ^ ((a ^ b) & 0x80808080u);
}

// XCD-aware workgroup order: the hardware deals workgroups round-robin over the 8 XCDs; this
// maps them so that each XCD receives a contiguous range of logical ids (shared operands stay in
// one L2). Bijective on [0, total).
__device__ __forceinline__ i64 xcd_remap(i64 h, i64 total) {
    const i64 nfull = total - (total & 7);
    if (h >= nfull) return h;
    return (h & 7) * (nfull >> 3) + (h >> 3);
}
