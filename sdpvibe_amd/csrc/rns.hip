// Exact multiprecision GEMM on CDNA4 INT8 matrix cores: residue split, modular products, CRT.
//
// Residues are centred, |r| <= (m-1)/2 <= 8128 (primes m <= 16257), and split r = 128*h + l with
// l in [-64, 63], h in [-64, 64]; then s = h + l also fits int8, so a residue product needs three
// INT8 products instead of four:  ra*rb = 16256*HH + 128*SS - 127*LL  (HH = ha.hb, SS = sa.sb,
// LL = la.lb).  All matrix work runs on v_mfma_i32_32x32x32_i8 (wave64, 16 bytes per lane).
//
// MFMA fragment mapping used throughout (32x32x32, one wave):
//   A operand: lane l supplies row (l % 32), bytes [16*(l/32), +16) of the 32-byte k slice;
//   B operand: lane l supplies column (l % 32), the same k bytes;
//   D:         lane l, register j holds (row 8*(j/4) + 4*(l/32) + j%4, column l % 32).
// A and B use the same k-slot assignment, so the dot product is exact whatever order the
// hardware pairs the bytes in.
#include <type_traits>

#include "common.hpp"

namespace {

struct PlaneSrc {
    const i32* p;
    i64 stride;
    __device__ __forceinline__ u32 operator()(int k) const { return (u32)p[(i64)k * stride]; }
};

struct LdsWords {
    const u32* p;
    __device__ __forceinline__ u32 operator()(int k) const { return p[k]; }
};

__device__ __forceinline__ v16i zero16() {
    v16i z;
#pragma unroll
    for (int j = 0; j < 16; j++) z[j] = 0;
    return z;
}
__device__ __forceinline__ v4i zero4() {
    v4i z;
#pragma unroll
    for (int j = 0; j < 4; j++) z[j] = 0;
    return z;
}
__device__ __forceinline__ v16i mfma(v4i a, v4i b, v16i c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// ================================================================== residue split
// One workgroup (4 waves) splits a tile of 16 rows x 16 consecutive k (one 16-byte chunk of every
// row of the residue planes).
// Phase 1: thread t takes one entry, with the lanes along the contiguous index (k when ldk == 1,
//   rows otherwise) so the limb loads coalesce; the limbs are staged in LDS with independent
//   loads (in the entry's own slot row when they fit there), shift-rounded to the row exponent into Lout words and written as bytes into the LDS
//   slot of the wave that owns the entry's row in phase 2 (wave w: rows 4w..4w+3, slot
//   (r % 4) * 16 + k): u - 128 for every byte but the top one, which goes in as the signed top
//   byte of the two's complement value, so the bytes carry the sign themselves.
// Phase 2, per wave and block of 32 primes: the byte dot products x . T_hi and x . T_lo run on MFMA
//   with the wave's 64 slots as rows and the primes as columns, so each lane holds four runs of 4
//   consecutive k of one row; r = 128 T_hi.x + T_lo.x + C1 is reduced to the centred
//   residue (nearest quotient in double, exact for |r| < 2^30) and split into 128 hi + lo. The two
//   lane halves swap 4-byte runs so that each lane holds whole 16-byte chunks (one row, one prime)
//   and stores them with 16-byte writes: 64 contiguous bytes per prime and wave.
struct LdsStrided {
    const u32* p;
    int stride;
    __device__ __forceinline__ u32 operator()(int k) const { return p[k * stride]; }
};

__device__ __forceinline__ u32 pack4(int a, int b, int c, int d) {  // low bytes of a, b, c, d
    const u32 ab = __builtin_amdgcn_perm((u32)b, (u32)a, 0x0c0c0400u);
    const u32 cd = __builtin_amdgcn_perm((u32)d, (u32)c, 0x0c0c0400u);
    return __builtin_amdgcn_perm(cd, ab, 0x05040100u);
}

// table chunks of 32 bytes: KSC = NBp / 32 (0: any, read from NBp); up to three chunks the kernel
// fits 128 VGPRs, four workgroups per CU
template <int KSC>
__global__ void __launch_bounds__(256, (KSC >= 1 && KSC <= 3) ? 4 : 1) k_rns_split(
    const i32* __restrict__ X, const i64* __restrict__ e, const i64* __restrict__ Erow, int L, int Lout, i64 plane,
    i64 nb, i64 rows, i64 ldb, i64 ldr, i64 ldk, i64 k0, i64 kc, i64 kp, const int8_t* __restrict__ Thi,
    const int8_t* __restrict__ Tlo, int NBp, const i32* __restrict__ C1,
    const i32* __restrict__ mods, const double* __restrict__ minv, int g0, int G, int8_t* __restrict__ Rh,
    int8_t* __restrict__ Rl, i32* __restrict__ flag, i64 ntiles_k, int stage) {
    SDPV_DYNAMIC_LDS(unsigned char, smem);
    // slot rows of XS bytes (multiple of 16); row S starts 16 * ((S / 16) % 4) bytes in, so that the
    // 16-row stride of phase 1 for operands with contiguous rows spreads over four bank groups
    const int XS = NBp + 64;
    const int RS = L + 1;                               // words per staged-limb row
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    unsigned char* xs = smem;                           // [256 slots][XS]
    // row tiles fastest: neighbouring workgroups write neighbouring 16-byte rows of every residue
    // chunk (one contiguous run per prime and chunk)
    const i64 ntr = (nb * rows + 15) / 16;
    const i64 tr = blockIdx.x % ntr, tk = blockIdx.x / ntr;
    const i64 row0 = tr * 16, kk0 = tk * 16;
    const i64 nrows = nb * rows;

    const int l32 = lane & 31, hsel = lane >> 5;
    const int KS = KSC ? KSC : NBp / 32, nPB = (G + 31) / 32;
    const i64 nck = kp >> 4;
    // prime-block operands: the first four 32-byte table chunks and the constants, loaded once per
    // block of primes and shared by both 32-slot halves
    struct PB {
        v4i fh[4], fl[4];
        int m, c1;
        double mi;
    };
    auto load_pb = [&](int pb, PB& o) {
        const int pr = pb * 32 + l32;
        const int g = g0 + (pr < G ? pr : 0);
        const int8_t* th = Thi + (i64)g * NBp + 16 * hsel;
        const int8_t* tl = Tlo + (i64)g * NBp + 16 * hsel;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = u < KS;
            o.fh[u] = ok ? *(const v4i*)(th + u * 32) : zero4();
            o.fl[u] = ok ? *(const v4i*)(tl + u * 32) : zero4();
        }
        o.m = mods[g];
        o.mi = minv[g];
        o.c1 = C1[g];
    };
    PB cur, nxt;
    load_pb(0, cur);                                    // first block's tables: in flight during phase 1

    // ---- phase 1: thread t -> entry (r, k) of the tile
    {
        const bool kmajor = ldk == 1;
        const int r = kmajor ? (t >> 4) : (t & 15), k = kmajor ? (t & 15) : (t >> 4);
        const int S = r * 16 + k;                       // phase-2 slot: wave S / 64, row S % 64
        u32* row = (u32*)(xs + (size_t)S * XS + ((S >> 4) & 3) * 16);
        // staged limbs: stage 1, a region of their own ([256 threads][RS] after the slots); stage 2,
        // the entry's own slot row, which the rounded words then overwrite in place (see below)
        u32* raw = stage == 2 ? row : (u32*)(xs + (size_t)256 * XS) + (size_t)t * RS;
        const i64 R = row0 + r, kg = kk0 + k;
        const bool valid = R < nrows && kg < kc;
        const i64 Rc = valid ? R : 0;
        const i64 off = valid ? (Rc / rows) * ldb + (Rc % rows) * ldr + (k0 + kg) * ldk : 0;
        const i64 ex = e ? e[off] : 0;                  // exponents first: their latency overlaps the limbs'
        const i64 er = e ? Erow[Rc] : 0;
        if (stage) {
            const i32* px = X + off;
            for (int w0 = 0; w0 < L; w0 += 32) {         // independent loads, up to 32 in flight
                u32 v[32];
#pragma unroll
                for (int u = 0; u < 32; u++) {
                    v[u] = (u32)*px;                    // words past L - 1 re-read word L - 1
                    if (w0 + u + 1 < L) px += plane;
                }
                if (stage == 2) {                       // 16-byte row: L rounded up to 4 words fits
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        if (w0 + 4 * q < L) {
                            v4i g4;
                            g4[0] = (int)v[4 * q]; g4[1] = (int)v[4 * q + 1];
                            g4[2] = (int)v[4 * q + 2]; g4[3] = (int)v[4 * q + 3];
                            *(v4i*)(raw + w0 + 4 * q) = g4;
                        }
                } else {
#pragma unroll
                    for (int u = 0; u < 32; u++)
                        if (w0 + u < L) raw[w0 + u] = v[u];
                }
            }
        }
        // words leave in groups of four (16-byte LDS writes: conflict-free at this row stride)
        u32 b0 = 0x80808080u, b1 = 0x80808080u, b2 = 0x80808080u, b3 = 0x80808080u;
        auto flush4 = [&](int w4) {
            v4i q4;
            q4[0] = (int)b0; q4[1] = (int)b1; q4[2] = (int)b2; q4[3] = (int)b3;
            *(v4i*)(row + 4 * w4) = q4;
            b0 = b1 = b2 = b3 = 0x80808080u;
        };
        auto put = [&](int wi, u32 v) {                 // bias every byte but the top one
            const u32 x = v ^ (wi == Lout - 1 ? 0x00808080u : 0x80808080u);
            switch (wi & 3) {
                case 0: b0 = x; break;
                case 1: b1 = x; break;
                case 2: b2 = x; break;
                default: b3 = x; flush4(wi >> 2);
            }
        };
        if (valid) {
            i64 s = 0;
            if (e) s = (ex == EZV) ? (i64)(1 << 30) : er - ex;
            // In place (stage 2), output word k is written after every read of source words <= k
            // when s >= 0 (the reads of a batch of 8 precede its writes, and later batches read
            // words >= k + 8). A left shift (s < 0) reads below k, so it re-reads the limbs from
            // global memory instead.
            const bool lds_src = stage == 1 || (stage == 2 && s >= 0);
            const int bad = lds_src ? shr_round_batched(LdsStrided{raw, 1}, L, s, Lout, put)
                                    : shr_round_batched(PlaneSrc{X + off, plane}, L, s, Lout, put);
            if (bad && flag) atomicOr(flag, 1);
        } else {
            for (int wi = 0; wi < Lout; wi++) put(wi, 0u);  // zero entry
        }
        if (Lout & 3) flush4(Lout >> 2);                  // last partial group (zero bytes beyond)
        for (int w4 = (Lout + 3) / 4; w4 < NBp / 16; w4++) flush4(w4);
    }
    __syncthreads();

    // ---- phase 2: wave w owns slots 64w .. 64w + 63 (rows 4w .. 4w + 3 of the tile)
    for (int pb = 0; pb < nPB; pb++) {
        if (pb + 1 < nPB) load_pb(pb + 1, nxt);        // next block's table chunks in flight
        const int pr = pb * 32 + l32;
        const bool pok = pr < G;
        const int g = g0 + (pok ? pr : 0);
        const int8_t* th = Thi + (i64)g * NBp + 16 * hsel;
        const int8_t* tl = Tlo + (i64)g * NBp + 16 * hsel;
        const int m = cur.m;
        const double mi = cur.mi, md = (double)m;
        const int c1 = cur.c1;
        // the two 32-slot blocks one after the other, against the same table chunks (two MFMA
        // chains each; one block's accumulators live at a time, so four waves fit a SIMD)
#pragma unroll
        for (int eb = 0; eb < 2; eb++) {
            v16i ah = zero16(), al = zero16();
            const int Sx = wave * 64 + eb * 32 + l32;   // slot row of this lane's fragment
            const unsigned char* xe = xs + (size_t)Sx * XS + ((Sx >> 4) & 3) * 16 + 16 * hsel;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (u < KS) {
                    const v4i fx = *(const v4i*)(xe + u * 32);
                    ah = mfma(fx, cur.fh[u], ah);
                    al = mfma(fx, cur.fl[u], al);
                }
            }
            for (int ks = 4; ks < KS; ks++) {            // wide operands (L > 32)
                const v4i fh = *(const v4i*)(th + ks * 32), fl = *(const v4i*)(tl + ks * 32);
                const v4i fx = *(const v4i*)(xe + ks * 32);
                ah = mfma(fx, fh, ah);
                al = mfma(fx, fl, al);
            }
            u32 dh[4], dl[4];                           // [jq]: 4 bytes at k = 8 (jq & 1) + 4 hsel of row 2 eb + jq / 2
#pragma unroll
            for (int jq = 0; jq < 4; jq++) {
                int hi[4], lo[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int j = 4 * jq + u;
                    const double x = (double)(ah[j] * 128 + al[j] + c1);
                    const int c = (int)fma(-rint(x * mi), md, x);        // centred residue
                    hi[u] = (c + 64) >> 7;
                    lo[u] = c - (hi[u] << 7);
                }
                dh[jq] = pack4(hi[0], hi[1], hi[2], hi[3]);
                dl[jq] = pack4(lo[0], lo[1], lo[2], lo[3]);
            }
            // swap runs across the halves: half 0 stores row 2 eb (jq 0, 1 of both halves), half 1
            // row 2 eb + 1 (jq 2, 3). Swapping (jq 0, jq 2) gives half 0 (own jq 0 = k 0-3, partner's
            // = k 4-7) and half 1 (partner's jq 2 = k 0-3, own = k 4-7); (jq 1, jq 3) likewise for
            // k 8-15.
            swap_halves(dh[0], dh[2]);
            swap_halves(dh[1], dh[3]);
            swap_halves(dl[0], dl[2]);
            swap_halves(dl[1], dl[3]);
            v4i vh, vl;
            vh[0] = (int)dh[0]; vh[1] = (int)dh[2]; vh[2] = (int)dh[1]; vh[3] = (int)dh[3];
            vl[0] = (int)dl[0]; vl[1] = (int)dl[2]; vl[2] = (int)dl[1]; vl[3] = (int)dl[3];
            const i64 R = row0 + 4 * wave + 2 * eb + hsel;
            if (pok && R < nrows) {
                const i64 o = (((i64)pr * nck + tk) * nrows + R) * 16;
                *(v4i*)(Rh + o) = vh;
                *(v4i*)(Rl + o) = vl;
            }
        }
        cur = nxt;
    }
}

// ================================================================== modular products
// Residues are stored chunk-major: byte k of row r of prime g at ((g*(kp/16) + k/16)*rows + r)*16
// + k%16, so 64 consecutive rows of one 16-byte k-chunk are 1 KB of contiguous memory.
// WM x WN waves, each BLM x BLN blocks of 32x32; BKB = 64 bytes (4 chunks) per stage, or 32 (2 chunks:
// smaller stages, more workgroups per CU). Stages are
// copied global -> LDS by direct loads (no registers), NST stages in flight, one barrier per stage.
// LDS stage layout: [A h | A l | B h | B l], each plane [chunk][tile row][16 B]: fragment reads of
// 32 consecutive rows are contiguous and conflict-free. The s = h + l operand is formed in
// registers (bytewise add) from the h and l fragments.
// STAGE: for a full tile whose rows are 16-byte aligned, each wave writes its reduced results
// through LDS (the stage buffers, free after the main loop) and stores 16 contiguous bytes per lane
// instead of one 2-byte entry per lane per row.
template <int WM, int WN, int BLM, int BLN, int NST, int P4, int BKB = 64, bool STAGE = false>
__global__ void __launch_bounds__(WM* WN * 64, BKB == 32 ? 4 : ((WM * WN <= 4) ? 2 : 1)) k_rns_gemm(
    const int8_t* __restrict__ Ah, const int8_t* __restrict__ Al, const int8_t* __restrict__ Bh,
    const int8_t* __restrict__ Bl, int G, int nb, i64 M, i64 N, i64 kp, int16_t* __restrict__ acc, i64 ld_g,
    const i32* __restrict__ mods, const double* __restrict__ minv, const i32* __restrict__ kcoef, int g0, int mode,
    i64 tiles_m, i64 tiles_n) {
    constexpr int NWAVE = WM * WN;
    constexpr int BM = WM * BLM * 32, BN = WN * BLN * 32;
    constexpr int PA = BKB * BM, PB = BKB * BN;     // bytes of one plane of one stage
    constexpr int SBYTES = 2 * PA + 2 * PB;
    constexpr int QA = BM * BKB / 1024, QB = BN * BKB / 1024;   // 1 KB direct loads per plane per stage
    constexpr int LPW = (2 * QA + 2 * QB) / NWAVE;  // direct loads per wave per stage
    static_assert((2 * QA + 2 * QB) % NWAVE == 0, "load split");
    SDPV_DYNAMIC_LDS(int8_t, smem8);

    const i64 total = tiles_m * tiles_n * (i64)G * nb;
    const i64 lg = xcd_remap(blockIdx.x, total);
    const i64 tn = lg % tiles_n;
    i64 rest = lg / tiles_n;
    const i64 tm = rest % tiles_m;
    rest /= tiles_m;
    const int gi = (int)(rest % G);
    const i64 b = rest / G;
    const bool syrk = (mode & SDPV_GEMM_SYRK_LOWER) != 0;
    if (syrk && tn > tm) return;

    const int t = threadIdx.x, lane = t & 63, l32 = lane & 31, hsel = lane >> 5;
    const int wave = __builtin_amdgcn_readfirstlane(t / 64);   // uniform: scalar load addressing
    const int wm = wave / WN, wn = wave % WN;
    const i64 nck = kp / 16;
    const i64 rowsA = (i64)nb * M, rowsB = (i64)nb * N;
    const i64 r0 = tm * BM, c0 = tn * BN;
    const i64 offA = ((i64)gi * nck * rowsA + b * M + r0) * 16;  // tile origin, chunk 0
    const i64 offB = ((i64)gi * nck * rowsB + b * N + c0) * 16;

    // direct loads of stage s into LDS buffer s % NST; rows past the matrix load its last row (kept
    // out of the results), so every stage issues exactly LPW loads per wave (the waits count them).
    // Per-lane byte offsets inside a stage are fixed (32-bit: the host keeps 64 rows x 16 B of a
    // plane below 2^32); the stage origin is uniform.
    const int rlimA = (int)(M - 1 - r0 < BM - 1 ? M - 1 - r0 : BM - 1);
    const int rlimB = (int)(N - 1 - c0 < BN - 1 ? N - 1 - c0 : BN - 1);
    u32 loff[LPW];
#pragma unroll
    for (int u = 0; u < LPW; u++) {
        const int q = wave * LPW + u;
        if (q < 2 * QA) {
            const int e = (q % QA) * 64 + lane;
            loff[u] = (u32)(((e / BM) * rowsA + min(e % BM, rlimA)) * 16);
        } else {
            const int e = ((q - 2 * QA) % QB) * 64 + lane;
            loff[u] = (u32)(((e / BN) * rowsB + min(e % BN, rlimB)) * 16);
        }
    }
    auto issue = [&](int s) {
        int8_t* sbase = smem8 + (s % NST) * SBYTES;
        const i64 kc0 = (i64)s * (BKB / 16);
#pragma unroll
        for (int u = 0; u < LPW; u++) {
            const int q = wave * LPW + u;                // uniform per wave
            if (q < 2 * QA) {
                const int p = q / QA, qq = q % QA;
                const int8_t* base = (p ? Al : Ah) + offA + kc0 * rowsA * 16;
                SDPV_LDS_DMA16(base + loff[u], sbase + p * PA + qq * 1024);
            } else {
                const int q2 = q - 2 * QA;
                const int p = q2 / QB, qq = q2 % QB;
                const int8_t* base = (p ? Bl : Bh) + offB + kc0 * rowsB * 16;
                SDPV_LDS_DMA16(base + loff[u], sbase + 2 * PA + p * PB + qq * 1024);
            }
        }
    };

    v16i cHH[BLM][BLN], cSS[BLM][BLN], cLL[BLM][BLN];
#pragma unroll
    for (int i = 0; i < BLM; i++)
#pragma unroll
        for (int j = 0; j < BLN; j++) {
            cHH[i][j] = zero16();
            cSS[i][j] = zero16();
            cLL[i][j] = zero16();
        }

    auto compute = [&](int s) {
        const int8_t* sA = smem8 + (s % NST) * SBYTES;
        const int8_t* sB = sA + 2 * PA;
#pragma unroll
        for (int ks = 0; ks < BKB / 32; ks++) {
            const int c = 2 * ks + hsel;
            v4i ah[BLM], al[BLM], as[BLM], bh[BLN], bl[BLN], bs[BLN];  // as/bs unused when P4
#pragma unroll
            for (int i = 0; i < BLM; i++) {
                const int off = (c * BM + (wm * BLM + i) * 32 + l32) * 16;
                ah[i] = *(const v4i*)(sA + off);
                al[i] = *(const v4i*)(sA + PA + off);
            }
#pragma unroll
            for (int j = 0; j < BLN; j++) {
                const int off = (c * BN + (wn * BLN + j) * 32 + l32) * 16;
                bh[j] = *(const v4i*)(sB + off);
                bl[j] = *(const v4i*)(sB + PB + off);
            }
            if (P4) {
                // four products, no bytewise adds: HH, LL and MID = HL + LH
#pragma unroll
                for (int i = 0; i < BLM; i++)
#pragma unroll
                    for (int j = 0; j < BLN; j++) {
                        cHH[i][j] = mfma(ah[i], bh[j], cHH[i][j]);
                        cLL[i][j] = mfma(al[i], bl[j], cLL[i][j]);
                        cSS[i][j] = mfma(ah[i], bl[j], cSS[i][j]);
                        cSS[i][j] = mfma(al[i], bh[j], cSS[i][j]);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < BLM; i++)
#pragma unroll
                    for (int q = 0; q < 4; q++) as[i][q] = (int)add_bytes((u32)ah[i][q], (u32)al[i][q]);
#pragma unroll
                for (int j = 0; j < BLN; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) bs[j][q] = (int)add_bytes((u32)bh[j][q], (u32)bl[j][q]);
#pragma unroll
                for (int i = 0; i < BLM; i++)
#pragma unroll
                    for (int j = 0; j < BLN; j++) {
                        cHH[i][j] = mfma(ah[i], bh[j], cHH[i][j]);
                        cLL[i][j] = mfma(al[i], bl[j], cLL[i][j]);
                        cSS[i][j] = mfma(as[i], bs[j], cSS[i][j]);
                    }
            }
        }
    };

    const int nst = (int)(kp / BKB);
#pragma unroll
    for (int s = 0; s < NST - 1; s++)
        if (s < nst) issue(s);
    // the MFMA loop runs at a raised wave priority, the epilogue at the normal one: a wave in its
    // epilogue yields issue slots to a workgroup still in its main loop on the same SIMD
    __builtin_amdgcn_s_setprio(1);
    for (int s = 0; s < nst; s++) {
        // this wave's loads of stage s are done once at most (stages issued after s) x LPW remain
        if (s + NST - 2 < nst) SDPV_WAIT_VMCNT((NST - 2) * LPW);
        else SDPV_WAIT_VMCNT(0);
        __syncthreads();
        if (s + NST - 1 < nst) issue(s + NST - 1);
        compute(s);
    }

    __builtin_amdgcn_s_setprio(0);
    // Epilogue: v = k0*HH + k1*SS + k2*LL (+ old) is exact in double (|v| < 2^48), reduced once.
    const int g = g0 + gi;
    const int m = mods[g];
    const double mi = minv[g], md = (double)m;
    // kcoef = (16256, 128, -127) * s mod m. Three products: ra rb = 16256 HH + 128 SS - 127 LL;
    // four: 16384 HH + 128 MID + LL, with 16384 = 16256 + 128 and 1 = 128 - 127.
    const double kc0 = (double)kcoef[3 * g], kc1 = (double)kcoef[3 * g + 1], kc2 = (double)kcoef[3 * g + 2];
    const double k0c = P4 ? kc0 + kc1 : kc0;
    const double k1c = kc1;
    const double k2c = P4 ? kc1 + kc2 : kc2;
    const bool accum = (mode & SDPV_GEMM_ACCUMULATE) != 0;
    const bool centred = (mode & SDPV_GEMM_CENTRED) != 0;
    int16_t* out = acc + (i64)g * ld_g + b * M * N;
    const bool full = r0 + BM <= M && c0 + BN <= N;  // uniform: no per-entry tests inside the matrix
    if (STAGE) {
        constexpr int SROW = 40;                     // int16 per staged row: 80 bytes, 16-byte aligned
        static_assert((size_t)WM * WN * BLM * BLN * 32 * SROW * 2 <= (size_t)NST * SBYTES, "staging area");
        __syncthreads();                             // every wave is done reading the stage buffers
        int16_t* sw = (int16_t*)smem8 + (size_t)wave * (BLM * BLN * 32 * SROW);
#pragma unroll
        for (int i = 0; i < BLM; i++)
#pragma unroll
            for (int j = 0; j < BLN; j++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int rb = 8 * (q >> 2) + 4 * hsel + (q & 3);
                    const double v = fma(k0c, (double)cHH[i][j][q], fma(k1c, (double)cSS[i][j][q], k2c * (double)cLL[i][j][q]));
                    // centred representative of the product part (the nearest quotient of the per-entry path)
                    sw[((i * BLN + j) * 32 + rb) * SROW + l32] = (int16_t)(int)fma(-rint(v * mi), md, v);
                }
        __syncthreads();
        // lane -> 8 consecutive columns of one row; (r + old) is reduced again when accumulating, which
        // gives the residue of the per-entry path's single reduction
        auto finish = [&](int r, const int16_t* po) {
            if (accum) {
                const double x = (double)r + (double)*po;
                r = (int)fma(-rint(x * mi), md, x);
            }
            if (!centred && r < 0) r += m;
            return r;
        };
        const bool vec = full && N % 8 == 0 && (((size_t)out) & 15) == 0;   // uniform
        const int seg = lane & 3;
#pragma unroll
        for (int i = 0; i < BLM; i++)
#pragma unroll
            for (int j = 0; j < BLN; j++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int rb = 16 * h + (lane >> 2);
                    const v4i w = *(const v4i*)(sw + ((i * BLN + j) * 32 + rb) * SROW + seg * 8);
                    const i64 row = r0 + (wm * BLM + i) * 32 + rb;
                    const i64 col = c0 + (wn * BLN + j) * 32 + seg * 8;
                    int16_t* o = out + row * N + col;
                    if (vec) {
                        v4i y = w;
                        if (accum || !centred) {
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const int lo = finish((int)(int16_t)(w[u] & 0xffff), o + 2 * u);
                                const int hi = finish(w[u] >> 16, o + 2 * u + 1);
                                y[u] = (int)(((u32)hi << 16) | ((u32)lo & 0xffffu));
                            }
                        }
                        *(v4i*)o = y;
                    } else if (row < M) {
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            if (col + u < N) {
                                const int r = (int)(int16_t)((u & 1) ? (w[u >> 1] >> 16) : (w[u >> 1] & 0xffff));
                                o[u] = (int16_t)finish(r, o + u);
                            }
                        }
                    }
                }
        return;
    }
#pragma unroll
    for (int i = 0; i < BLM; i++)
#pragma unroll
        for (int j = 0; j < BLN; j++) {
            const i64 col = c0 + (wn * BLN + j) * 32 + l32;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const i64 row = r0 + (wm * BLM + i) * 32 + 8 * (q >> 2) + 4 * hsel + (q & 3);
                if (full || (row < M && col < N)) {
                    double v = fma(k0c, (double)cHH[i][j][q], fma(k1c, (double)cSS[i][j][q], k2c * (double)cLL[i][j][q]));
                    if (accum) v += (double)out[row * N + col];
                    // nearest quotient, exact: v*mi errs by < 2^-17, v/m is >= 1/(2m) > 2^-15 from a half-integer
                    int r = (int)fma(-rint(v * mi), md, v);
                    if (!centred && r < 0) r += m;
                    out[row * N + col] = (int16_t)r;
                }
            }
        }
}

template <int WM, int WN, int BLM, int BLN, int NST, int P4, int BKB = 64, bool STAGE = false>
int launch_gemm(const int8_t* Ah, const int8_t* Al, const int8_t* Bh, const int8_t* Bl, int G, int nb, i64 M, i64 N,
                i64 kp, int16_t* acc, i64 ld_g, const i32* mods, const double* minv, const i32* kcoef, int g0,
                int mode, hipStream_t st) {
    constexpr int BM = WM * BLM * 32, BN = WN * BLN * 32;
    const i64 tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    const i64 total = tiles_m * tiles_n * (i64)G * nb;
    if (total == 0) return SDPV_OK;
    if (total > 0x7fffffffLL) return SDPV_ESIZE;
    const size_t lds = (size_t)NST * 2 * BKB * (BM + BN);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)k_rns_gemm<WM, WN, BLM, BLN, NST, P4, BKB, STAGE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr_set = true;
    }
    hipLaunchKernelGGL((k_rns_gemm<WM, WN, BLM, BLN, NST, P4, BKB, STAGE>), dim3((unsigned)total), dim3(WM * WN * 64), lds, st, Ah,
                       Al, Bh, Bl, G, nb, M, N, kp, acc, ld_g, mods, minv, kcoef, g0, mode, tiles_m, tiles_n);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

// ================================================================== CRT reconstruction
// One wavefront (one workgroup) reconstructs 64 entries; lane l owns entry q0 + l.
// Phase 1 (lane = entry, every prime; the prime index is uniform, so its constants and plane
//   address are scalar): y_g = centred(acc_g * inv_g mod m_g), or acc_g itself when the product
//   stored it so (SCALED); split to 128*hi + lo and written as packed bytes to the lane's LDS rows;
//   k = round(sum y_g/m_g) (|sum - k| < 1/4 because |value| < M/4), and -k into column n, whose
//   Ct column holds the digits of M.
// Phase 2, per block of 32 base-256 digits and per group of 32 entries: Z = Ct . Y on MFMA with
//   digits as rows and entries as columns (Z_d = digit d of sum_g y_g M/m_g - k M). Lanes 0-31 then
//   hold the digits of words 0, 2, 4, 6 and lanes 32-63 those of words 1, 3, 5, 7 of the group's
//   entries; each lane folds its digits into word partials sum Z_d 256^u, one lane-half swap gives
//   every lane the 8 partials of its own entry, and it carries them into words.
// Phase 3 (lane = entry): words -> W + bitlen (mode 0), normalised MPF limbs (1), or a
//   fixed-exponent rounding (2).

// NDB > 0 (modes 1, 2 with that many digit blocks): the words of an entry stay in registers
// during phase 2 and move into the then unused byte rows afterwards, so a wave needs only the
// byte rows of LDS (more waves per CU); NDB = 0: words in their own LDS rows, digit blocks
// pipelined.
template <bool SCALED, int NDB>
__global__ void __launch_bounds__(64) k_rns_crt(
    const int16_t* __restrict__ acc, int n, i64 ld_g, i64 rows, i64 N, i64 j0, i64 w, const i32* __restrict__ mods,
    const double* __restrict__ minv, const i32* __restrict__ inv, const int8_t* __restrict__ Ct, int E, int npad,
    int mode, i32* __restrict__ W, i64* __restrict__ bitlen, i32* __restrict__ outm, i64* __restrict__ oute, int L,
    i64 plane, i64 ldo, i64 jo, i64 Mb, const i64* __restrict__ erow, const i64* __restrict__ ecol,
    const i64* __restrict__ scol, i32* __restrict__ flag) {
    SDPV_DYNAMIC_LDS(unsigned char, smem);
    const int YS = npad + 16;                         // bytes per entry row (multiple of 16)
    const int NW = E / 4;
    const int Epad = ((E + 31) / 32) * 32;
    const int lane = threadIdx.x, l32 = lane & 31, half = lane >> 5;
    int8_t* Yh = (int8_t*)smem;                       // [64][YS]
    int8_t* Yl = Yh + 64 * YS;                        // [64][YS]
    u32* wrow = (u32*)(NDB > 0 ? (unsigned char*)smem : (unsigned char*)(Yl + 64 * YS)) + (size_t)lane * (NW + 1);
                                                      // [64][NW + 1] words (modes 1, 2)

    const i64 Bq = rows * w;
    const i64 q = (i64)blockIdx.x * 64 + lane;
    const bool qok = q < Bq;
    // byte offset of the entry in a prime plane (< 2^32: the host checks ld_g < 2^31); lanes past
    // the block read entry 0 and store nothing
    const u32 boff = qok ? (u32)(((q / w) * N + j0 + q % w) * 2) : 0u;

    // ---- phase 1
    double sp[4] = {0.0, 0.0, 0.0, 0.0};             // four independent partial sums of y_g / m_g
    int8_t* yhr = Yh + (size_t)lane * YS;
    int8_t* ylr = Yl + (size_t)lane * YS;
    // 16 primes gc .. gc + 15; FULL: all below n (consecutive constants: wide scalar loads)
    auto chunk = [&](int gc, auto full_t) {
        constexpr bool FULL = decltype(full_t)::value;
        const int cnt = n - gc;                       // primes in this chunk (uniform)
        int a[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
            a[u] = (FULL || u < cnt) ? *(const int16_t*)((const char*)(acc + (i64)(gc + u) * ld_g) + boff) : 0;
        int hi[16], lo[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int g = FULL ? gc + u : (gc + u < n ? gc + u : n - 1);  // padding primes: a = 0, y = 0
            int y = a[u];
            if (!SCALED) {                              // x < m^2 < 2^28: exact nearest quotient
                const double x = (double)y * (double)inv[g];
                y = (int)fma(-rint(x * minv[g]), (double)mods[g], x);
            }
            sp[u & 3] = fma((double)y, minv[g], sp[u & 3]);
            hi[u] = (y + 64) >> 7;
            lo[u] = y - (hi[u] << 7);
        }
        v4i vh, vl;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            vh[c] = (int)pack4(hi[4 * c], hi[4 * c + 1], hi[4 * c + 2], hi[4 * c + 3]);
            vl[c] = (int)pack4(lo[4 * c], lo[4 * c + 1], lo[4 * c + 2], lo[4 * c + 3]);
        }
        *(v4i*)(yhr + gc) = vh;
        *(v4i*)(ylr + gc) = vl;
    };
    int gc = 0;
    for (; gc + 16 <= n; gc += 16) chunk(gc, std::true_type{});
    for (; gc < npad; gc += 16) chunk(gc, std::false_type{});
    {                                                 // column n: -k = 128*hi + lo (|k| <= n/2)
        const int nk = -(int)rint((sp[0] + sp[1]) + (sp[2] + sp[3]));
        const int kl = ((nk + 64) & 127) - 64;
        yhr[n] = (int8_t)((nk - kl) >> 7);
        ylr[n] = (int8_t)kl;
    }
    wave_sync();

    // ---- phase 2
    const int nDB = Epad / 32, nK = npad / 32;
    const int8_t* y0h = Yh + (size_t)l32 * YS + 16 * half;   // B operand of entry group 0
    const int8_t* y0l = Yl + (size_t)l32 * YS + 16 * half;
    const int8_t* y1h = y0h + 32 * YS;                        // group 1
    const int8_t* y1l = y0l + 32 * YS;
    i64 carry = 0;
    int last0 = -1, last1 = -1;
    u32 w0 = 0, w1 = 0, lastw = 0;
    // the first KF 32-byte digit-table chunks of a digit block live in registers. Software pipeline
    // over digit blocks: the MFMAs of block d + 1 run while block d is folded and carried, and the
    // table chunks of block d + 2 load meanwhile.
    constexpr int KF = 4;
    auto load_chunks = [&](int db, v4i* cf) {
        const int8_t* crow = Ct + (i64)(db * 32 + l32) * npad + 16 * half;
#pragma unroll
        for (int u = 0; u < KF; u++) cf[u] = u < nK ? *(const v4i*)(crow + u * 32) : zero4();
    };
    auto contract = [&](int db, const v4i* cf, v16i* z) {   // z: zh0, zl0, zh1, zl1
        z[0] = mfma(cf[0], *(const v4i*)y0h, zero16());
        z[1] = mfma(cf[0], *(const v4i*)y0l, zero16());
        z[2] = mfma(cf[0], *(const v4i*)y1h, zero16());
        z[3] = mfma(cf[0], *(const v4i*)y1l, zero16());
#pragma unroll
        for (int u = 1; u < KF; u++) {
            if (u < nK) {
                const int o = u * 32;
                z[0] = mfma(cf[u], *(const v4i*)(y0h + o), z[0]);
                z[1] = mfma(cf[u], *(const v4i*)(y0l + o), z[1]);
                z[2] = mfma(cf[u], *(const v4i*)(y1h + o), z[2]);
                z[3] = mfma(cf[u], *(const v4i*)(y1l + o), z[3]);
            }
        }
        const int8_t* crow = Ct + (i64)(db * 32 + l32) * npad + 16 * half;
        for (int kb = KF; kb < nK; kb++) {           // wide prime sets (more than 32 KF primes)
            const v4i c = *(const v4i*)(crow + kb * 32);
            const int o = kb * 32;
            z[0] = mfma(c, *(const v4i*)(y0h + o), z[0]);
            z[1] = mfma(c, *(const v4i*)(y0l + o), z[1]);
            z[2] = mfma(c, *(const v4i*)(y1h + o), z[2]);
            z[3] = mfma(c, *(const v4i*)(y1l + o), z[3]);
        }
    };
    auto fold = [&](int db, const v16i* z, u32* word) {
        // word partials sum_u Z_u 256^u, Z_u = 128 zh + zl (|Z_u| < 2^31); X: group 0, Y: group 1
        u32 xl[4], xh[4], yl[4], yh[4];
#pragma unroll
        for (int wq = 0; wq < 4; wq++) {
            i64 v0 = 0, v1 = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                v0 += (i64)(128 * z[0][4 * wq + u] + z[1][4 * wq + u]) << (8 * u);
                v1 += (i64)(128 * z[2][4 * wq + u] + z[3][4 * wq + u]) << (8 * u);
            }
            xl[wq] = (u32)v0;
            xh[wq] = (u32)((u64)v0 >> 32);
            yl[wq] = (u32)v1;
            yh[wq] = (u32)((u64)v1 >> 32);
            // lanes 0-31 of X hold the even words of group 0's entries (their owners), lanes 32-63
            // of X the odd ones; Y likewise for group 1, owned by lanes 32-63. After the swap, X
            // holds the even and Y the odd word partial of the lane's own entry.
            swap_halves(xl[wq], yl[wq]);
            swap_halves(xh[wq], yh[wq]);
        }
#pragma unroll
        for (int h2 = 0; h2 < 8; h2++) {
            const int wq = h2 >> 1;
            const u64 p = (h2 & 1) ? ((u64)yh[wq] << 32 | yl[wq]) : ((u64)xh[wq] << 32 | xl[wq]);
            const i64 v = carry + (i64)p;
            word[h2] = (u32)v;
            carry = v >> 32;
        }
    };
    auto store_words = [&](int db, const u32* word) {
        const int nw = NW - db * 8;                    // words of this block inside the value
        if (mode == SDPV_CRT_WORDS) {
#pragma unroll
            for (int h2 = 0; h2 < 8; h2++) {
                if (h2 < nw) {
                    const int wi = db * 8 + h2;
                    if (qok) W[(i64)wi * Bq + q] = (i32)word[h2];
                    if (word[h2] != 0u) { last0 = wi; w0 = word[h2]; }
                    if (word[h2] != 0xffffffffu) { last1 = wi; w1 = ~word[h2]; }
                    lastw = word[h2];
                }
            }
        } else {
#pragma unroll
            for (int h2 = 0; h2 < 8; h2++)
                if (h2 < nw) wrow[db * 8 + h2] = word[h2];
        }
    };
    if constexpr (NDB > 0) {
        u32 wv[NDB * 8];                                // the entry's words (static indices)
        v4i cfc[KF], cfn[KF];
        load_chunks(0, cfc);
#pragma unroll
        for (int db = 0; db < NDB; db++) {
            if (db + 1 < NDB) load_chunks(db + 1, cfn);
            v16i z[4];
            contract(db, cfc, z);
            fold(db, z, wv + 8 * db);
#pragma unroll
            for (int u = 0; u < KF; u++) cfc[u] = cfn[u];
        }
        wave_sync();                                    // the byte rows are read: reuse them
#pragma unroll
        for (int i = 0; i < NDB * 8; i++)
            if (i < NW) wrow[i] = wv[i];
    } else {
        v4i cfA[KF], cfB[KF];
        v16i zA[4], zB[4];
        u32 word[8];
        load_chunks(0, cfA);
        contract(0, cfA, zA);
        if (nDB > 1) load_chunks(1, cfB);
        for (int db = 0; db < nDB; db += 2) {
            if (db + 1 < nDB) {
                contract(db + 1, cfB, zB);
                if (db + 2 < nDB) load_chunks(db + 2, cfA);
            }
            fold(db, zA, word);
            store_words(db, word);
            if (db + 1 < nDB) {
                if (db + 2 < nDB) {
                    contract(db + 2, cfA, zA);
                    if (db + 3 < nDB) load_chunks(db + 3, cfB);
                }
                fold(db + 1, zB, word);
                store_words(db + 1, word);
            }
        }
    }

    // ---- phase 3 (lane = entry; reads only its own LDS row, in batches of independent reads)
    if (!qok) return;
    const i64 i = q / w, jj = q % w;
    if (mode == SDPV_CRT_WORDS) {
        const bool neg = (lastw >> 31) != 0;
        const int last = neg ? last1 : last0;
        const u32 top = neg ? w1 : w0;
        bitlen[q] = last >= 0 ? 32 * (i64)last + 32 - clz32(top) : 0;
        return;
    }
    const i64 o = i * ldo + jo + jj;
    const i64 cidx = (i / Mb) * ldo + jo + jj;
    const u32 ext = (wrow[NW - 1] >> 31) ? 0xffffffffu : 0u;   // sign extension word
    const LdsWords src{wrow};
    // word t of the value's two's complement: zero below 0, the sign extension above NW - 1
    auto wd = [&](int t) -> u32 {
        const u32 v = wrow[t < 0 ? 0 : (t >= NW ? NW - 1 : t)];
        return t < 0 ? 0u : (t >= NW ? ext : v);
    };
    i64 sv;                                           // shift: out = (v + rounding) / 2^sv
    u32 c = 0;                                        // rounding carry into bit sv
    if (mode == SDPV_CRT_NORMALIZE) {
        // bit length of |v|: top word T of v ^ ext; for v < 0 also the lowest nonzero word z, as
        // |v| = ~v above z, -v_z at z (mpf.py normalize via bitlen of the magnitude)
        const bool neg = ext != 0u;
        int T = -1;
        u32 wT = 0;
        for (int b = NW - 1; b >= 0 && T < 0; b -= 8) {
            u32 x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = (b - u >= 0) ? wd(b - u) ^ ext : 0u;
#pragma unroll
            for (int u = 7; u >= 0; u--)
                if (x[u] != 0u) { T = b - u; wT = x[u]; }
        }
        int z = -1;
        u32 vz = 0;
        if (neg) {
            for (int b = 0; b < NW && z < 0; b += 8) {
                u32 x[8];
#pragma unroll
                for (int u = 0; u < 8; u++) x[u] = wd(b + u);
#pragma unroll
                for (int u = 7; u >= 0; u--)
                    if (x[u] != 0u && b + u < NW) { z = b + u; vz = x[u]; }
            }
        }
        int bl;
        if (!neg) bl = T < 0 ? 0 : 32 * T + 32 - clz32(wT);
        else bl = T > z ? 32 * T + 32 - clz32(wT) : 32 * z + 32 - clz32(0u - vz);
        if (bl == 0) {
            for (int k = 0; k < L; k++) outm[(i64)k * plane + o] = 0;
            oute[o] = EZV;
            return;
        }
        sv = bl - (32 * L - 2);
        if (sv > 0) {
            // -round_half_up(|v| / 2^sv) = floor((v + 2^(sv-1) - 1) / 2^sv) for v < 0: the carry into
            // bit sv needs bit sv-1 and, for v < 0, a nonzero bit below it
            const u32 rb = (wd((int)((sv - 1) >> 5)) >> (int)((sv - 1) & 31)) & 1u;
            const bool sticky = !neg || 32 * (i64)z + __builtin_ctz(vz) < sv - 1;
            c = (rb && sticky) ? 1u : 0u;
        }
        const i64 ein = (erow ? erow[i] : 0) + (ecol ? ecol[cidx] : 0);
        oute[o] = ein + sv;
        (void)shift_words(src, NW, ext, sv, c, L, [&](int k, u32 v) { outm[(i64)k * plane + o] = (i32)v; });
    } else {
        // round_half_up(v / 2^s) = floor((v + 2^(s-1)) / 2^s) (limbops._shift_round_kernel)
        const int bad = shr_round_batched(src, NW, scol[cidx], L, [&](int k, u32 v) { outm[(i64)k * plane + o] = (i32)v; });
        if (bad && flag) atomicOr(flag, 1);
    }
}

// ================================================================== MFMA layout self-test
__global__ void __launch_bounds__(64) k_mfma_selftest(const int8_t* A, const int8_t* B, i32* D) {
    const int lane = threadIdx.x, l32 = lane & 31, hsel = lane >> 5;
    v4i fa, fb;
    int8_t* pa = (int8_t*)&fa;
    int8_t* pb = (int8_t*)&fb;
    for (int u = 0; u < 16; u++) {
        pa[u] = A[l32 * 32 + 16 * hsel + u];
        pb[u] = B[(16 * hsel + u) * 32 + l32];
    }
    v16i d = mfma(fa, fb, zero16());
    for (int j = 0; j < 16; j++) D[(8 * (j >> 2) + 4 * hsel + (j & 3)) * 32 + l32] = d[j];
}

inline unsigned nblk(i64 n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace

extern "C" {

int sdpv_rns_split(const int32_t* X, const int64_t* e, const int64_t* Erow, int64_t L, int64_t Lout, int64_t plane,
                   int64_t nb, int64_t rows, int64_t ldb, int64_t ldr, int64_t ldk, int64_t k0, int64_t kc,
                   int64_t kp, const int8_t* Thi, const int8_t* Tlo, int64_t NBp, const int32_t* C1,
                   const int32_t* mods, const double* minv, int64_t g0, int64_t G, int8_t* Rh,
                   int8_t* Rl, int32_t* flag, void* stream) {
    if (L < 1 || Lout < 1 || nb < 0 || rows < 0 || kc < 0 || kp < kc || G < 0 || g0 < 0) return SDPV_EARG;
    if (kp % 64 != 0 || NBp != 32 * ((4 * Lout + 31) / 32)) return SDPV_EARG;
    if (Lout > 128 || L > 128) return SDPV_ESIZE;
    const i64 nrows = nb * rows;
    if (nrows == 0 || kp == 0 || G == 0) return SDPV_OK;
    if (!X || !Thi || !Tlo || !C1 || !mods || !minv || !Rh || !Rl) return SDPV_EARG;
    if (e && !Erow) return SDPV_EARG;
    const i64 ntiles_k = kp / 16, ntiles_r = (nrows + 15) / 16;
    const i64 grid = ntiles_r * ntiles_k;
    if (grid > 0x7fffffffLL) return SDPV_ESIZE;
    // the limbs are staged in LDS: in the slot rows themselves when L words fit in one after its
    // offset of up to 48 bytes (stage 2: no extra LDS, 40 KB per workgroup at L = 24, so four
    // workgroups fit a CU), else after them when that fits (1)
    size_t lds = (size_t)256 * (NBp + 64);
    const size_t stage_bytes = (size_t)256 * 4 * (L + 1);
    int stage = 0;
    if (NBp + 16 >= 4 * L) stage = 2;
    else if (lds + stage_bytes <= 160 * 1024) stage = 1;
    if (stage == 1) lds += stage_bytes;
    if (lds > 160 * 1024) return SDPV_ESIZE;
    static bool attr_set[5] = {false, false, false, false, false};
    const int ks = NBp / 32 <= 4 ? (int)(NBp / 32) : 0;
#define SDPV_SPLIT_LAUNCH(KSC)                                                                                 \
    do {                                                                                                       \
        if (!attr_set[KSC]) {                                                                                  \
            (void)hipFuncSetAttribute((const void*)k_rns_split<KSC>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                      160 * 1024);                                                             \
            attr_set[KSC] = true;                                                                              \
        }                                                                                                      \
        hipLaunchKernelGGL((k_rns_split<KSC>), dim3((unsigned)grid), dim3(256), lds, as_stream(stream), X, e,   \
                           Erow, (int)L, (int)Lout, (i64)plane, (i64)nb, (i64)rows, (i64)ldb, (i64)ldr, (i64)ldk, \
                           (i64)k0, (i64)kc, (i64)kp, Thi, Tlo, (int)NBp, C1, mods, minv, (int)g0, (int)G, Rh, Rl, \
                           flag, ntiles_k, stage);                                                             \
    } while (0)
    switch (ks) {
        case 1: SDPV_SPLIT_LAUNCH(1); break;
        case 2: SDPV_SPLIT_LAUNCH(2); break;
        case 3: SDPV_SPLIT_LAUNCH(3); break;
        case 4: SDPV_SPLIT_LAUNCH(4); break;
        default: SDPV_SPLIT_LAUNCH(0); break;
    }
#undef SDPV_SPLIT_LAUNCH
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_rns_gemm(const int8_t* Ah, const int8_t* Al, const int8_t* Bh, const int8_t* Bl, int64_t G, int64_t nb,
                  int64_t M, int64_t N, int64_t kp, int16_t* acc, int64_t ld_g, const int32_t* mods,
                  const double* minv, const int32_t* kcoef, int64_t g0, int32_t mode, int32_t config, void* stream) {
    if (G < 0 || nb < 0 || M < 0 || N < 0 || kp < 0 || g0 < 0 || (mode & ~7)) return SDPV_EARG;
    if (config < -1 || config > 13) return SDPV_EARG;
    if (kp % 64 != 0 || kp > 65536) return SDPV_ESIZE;
    if (nb * M > (1LL << 26) || nb * N > (1LL << 26)) return SDPV_ESIZE;   // 32-bit offsets in a stage
    // the syrk mode skips tiles above the diagonal by tile index: square tiles only
    if ((mode & SDPV_GEMM_SYRK_LOWER) &&
        (M != N || !(config == -1 || config == 0 || config == 1 || config == 2 || config == 3 || config == 6)))
        return SDPV_EARG;
    if (G == 0 || nb == 0 || M == 0 || N == 0) return SDPV_OK;
    if (!Ah || !Al || !Bh || !Bl || !acc || !mods || !minv || !kcoef) return SDPV_EARG;
    hipStream_t st = as_stream(stream);
    if (kp == 0) return SDPV_EARG;
    if (config == -1) {
        const i64 mn = M < N ? M : N;  // syrk: square tiles of 128 / 64 / 32 by N, as the reference
        config = mn >= 128 ? 6 : (mn >= 64 ? 2 : 3);   // 128x128 four-product tile: the fastest
    }
    const int g = (int)G, b = (int)nb, q0 = (int)g0;
    switch (config) {
        case 0: return launch_gemm<2, 2, 2, 2, 2, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 1: return launch_gemm<2, 2, 2, 2, 3, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 2: return launch_gemm<2, 2, 1, 1, 2, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 3: return launch_gemm<1, 1, 1, 1, 2, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 4: return launch_gemm<4, 2, 2, 2, 2, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 5: return launch_gemm<2, 4, 2, 2, 2, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 6: return launch_gemm<2, 2, 2, 2, 2, 1>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 7: return launch_gemm<2, 4, 2, 2, 2, 1>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 8: return launch_gemm<2, 4, 2, 2, 3, 1>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 9: return launch_gemm<2, 4, 2, 2, 3, 0>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 10: return launch_gemm<2, 2, 2, 1, 2, 1>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 11: return launch_gemm<2, 2, 1, 2, 2, 1>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        case 12: return launch_gemm<2, 2, 1, 2, 3, 1, 32>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
        // configuration 12 with the epilogue staged through LDS (16-byte stores)
        default: return launch_gemm<2, 2, 1, 2, 3, 1, 32, true>(Ah, Al, Bh, Bl, g, b, M, N, kp, acc, ld_g, mods, minv, kcoef, q0, mode, st);
    }
}

int sdpv_rns_crt(const int16_t* acc, int64_t n, int64_t ld_g, int64_t rows, int64_t N, int64_t j0, int64_t w,
                 const int32_t* mods, const double* minv, const int32_t* inv, const int8_t* Ct, int64_t E,
                 int64_t npad, int32_t mode, int32_t* W, int64_t* bitlen, int32_t* outm, int64_t* oute, int64_t L,
                 int64_t plane, int64_t ldo, int64_t jo, int64_t Mb, const int64_t* erow, const int64_t* ecol,
                 const int64_t* scol, int32_t* flag, void* stream) {
    if (n < 1 || rows < 0 || N < 0 || j0 < 0 || w < 0 || E < 4 || E % 4 != 0 || npad <= n || npad % 32 != 0)
        return SDPV_EARG;
    const int omode = mode & ~SDPV_CRT_SCALED;
    if (omode < 0 || omode > 2) return SDPV_EARG;
    const i64 Bq = rows * w;
    if (Bq == 0) return SDPV_OK;
    if (!acc || !mods || !minv || !inv || !Ct) return SDPV_EARG;
    if (omode == SDPV_CRT_WORDS && (!W || !bitlen)) return SDPV_EARG;
    if (omode != SDPV_CRT_WORDS && (!outm || L < 1 || Mb < 1)) return SDPV_EARG;
    if (omode == SDPV_CRT_NORMALIZE && !oute) return SDPV_EARG;
    if (omode == SDPV_CRT_ROUND && !scol) return SDPV_EARG;
    if (npad > 1024) return SDPV_ESIZE;  // digit sums 128*Zh + Zl stay within int32
    if ((rows - 1) * N + j0 + w > ld_g || ld_g >= (1LL << 31)) return SDPV_ESIZE;  // 32-bit plane offsets
    const i64 NW = E / 4;
    const int ndb = (int)((E + 31) / 32);
    const bool scaled = (mode & SDPV_CRT_SCALED) != 0;
    // register-held words: scaled input, limb output, 5..8 digit blocks (about L = 16..31; more
    // words would cost the wave count that the smaller LDS gains)
    const int reg = (scaled && omode != SDPV_CRT_WORDS && ndb >= 5 && ndb <= 8) ? ndb : 0;
    const size_t ybytes = (size_t)128 * (npad + 16), wbytes = (size_t)256 * (NW + 1);
    const size_t lds = omode == SDPV_CRT_WORDS ? ybytes : (reg ? (ybytes > wbytes ? ybytes : wbytes) : ybytes + wbytes);
    if (lds > 160 * 1024) return SDPV_ESIZE;
    const i64 grid = (Bq + 63) / 64;
    if (grid > 0x7fffffffLL) return SDPV_ESIZE;
    hipStream_t st = as_stream(stream);
#define SDPV_CRT_LAUNCH(SC, ND)                                                                                \
    do {                                                                                                       \
        static bool attr = false;                                                                              \
        if (!attr) {                                                                                           \
            (void)hipFuncSetAttribute((const void*)k_rns_crt<SC, ND>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                      160 * 1024);                                                             \
            attr = true;                                                                                       \
        }                                                                                                      \
        hipLaunchKernelGGL((k_rns_crt<SC, ND>), dim3((unsigned)grid), dim3(64), lds, st, acc, (int)n, (i64)ld_g, \
                           (i64)rows, (i64)N, (i64)j0, (i64)w, mods, minv, inv, Ct, (int)E, (int)npad, omode, W,  \
                           bitlen, outm, oute, (int)L, (i64)plane, (i64)ldo, (i64)jo, (i64)Mb, erow, ecol, scol,  \
                           flag);                                                                              \
    } while (0)
    if (!scaled) SDPV_CRT_LAUNCH(false, 0);
    else switch (reg) {
        case 5: SDPV_CRT_LAUNCH(true, 5); break;
        case 6: SDPV_CRT_LAUNCH(true, 6); break;
        case 7: SDPV_CRT_LAUNCH(true, 7); break;
        case 8: SDPV_CRT_LAUNCH(true, 8); break;
        default: SDPV_CRT_LAUNCH(true, 0); break;
    }
#undef SDPV_CRT_LAUNCH
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

int sdpv_mfma_selftest(const int8_t* A, const int8_t* B, int32_t* D, void* stream) {
    if (!A || !B || !D) return SDPV_EARG;
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, as_stream(stream), A, B, D);
    SDPV_TRY_LAUNCH();
    return SDPV_OK;
}

}  // extern "C"
