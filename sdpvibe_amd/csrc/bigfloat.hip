// SDPB binary block data <-> MPF entries (host code; no device work).
//
// SDPB's pmp2sdp writes El::BigFloat values through boost::archive::binary_oarchive. One record is
//     uint8 is_zero; if !is_zero: int32 size, int64 exp, nl x uint64 limbs (a GMP mpf_t),
// value = sign(size) * D * 2^(64 (exp - |size|)), D = sum_{i < |size|} limb_i 2^(64 i); limbs from
// |size| to nl - 1 carry no meaning (sdpvibe/_bigfloat.c:1-11 states the layout).
//
// Decoding rounds D to at most NB = 32L - 2 bits (magnitude half up, as every MPF rounding here) and
// stores it as L 32-bit two's-complement limbs, entry-major, with a binary exponent; encoding writes
// the exact value back with the smallest number of 64-bit limbs.
#include <cstring>

#include "common.hpp"

namespace {

// bit length of the n-limb magnitude d (0 for zero)
int64_t bit_length(const uint64_t* d, int64_t n) {
    for (int64_t i = n - 1; i >= 0; i--)
        if (d[i]) return 64 * i + 64 - __builtin_clzll(d[i]);
    return 0;
}

// bit b of the little-endian limb array d (n limbs; 0 beyond)
inline unsigned bit_at(const uint64_t* d, int64_t n, int64_t b) {
    const int64_t w = b >> 6;
    return w < n ? (unsigned)((d[w] >> (b & 63)) & 1u) : 0u;
}

// 64 bits of d starting at bit position b
inline uint64_t bits_from(const uint64_t* d, int64_t n, int64_t b) {
    const int64_t w = b >> 6, s = b & 63;
    const uint64_t lo = w < n ? d[w] : 0, hi = (w + 1) < n ? d[w + 1] : 0;
    return s ? (lo >> s) | (hi << (64 - s)) : lo;
}

}  // namespace

extern "C" {

int64_t sdpv_decode_bigfloats(const uint8_t* buf, int64_t len, int64_t pos, int64_t n, int32_t nl, int32_t L,
                              int32_t* out, int64_t* exps, int64_t ez) {
    if (!buf || pos < 0 || n < 0 || nl < 1 || L < 1 || (n && (!out || !exps))) return -1;
    const int64_t NB = 32 * (int64_t)L - 2;
    const int64_t rl = (L + 1) / 2 + 1;                    // 64-bit limbs of the rounded magnitude
    uint64_t* d = new uint64_t[nl];
    uint64_t* r = new uint64_t[rl];
    int64_t result = -1;
    for (int64_t i = 0; i < n; i++) {
        int32_t* o = out + i * (int64_t)L;
        if (pos >= len) goto done;
        if (buf[pos++]) {                                  // is_zero
            std::memset(o, 0, 4 * (size_t)L);
            exps[i] = ez;
            continue;
        }
        if (pos + 12 + 8 * (int64_t)nl > len) goto done;
        int32_t size;
        int64_t ex;
        std::memcpy(&size, buf + pos, 4);
        std::memcpy(&ex, buf + pos + 4, 8);
        pos += 12;
        const int64_t asz = size < 0 ? -(int64_t)size : size;
        if (asz > nl) goto done;
        std::memcpy(d, buf + pos, 8 * (size_t)asz);
        pos += 8 * (int64_t)nl;
        const int64_t bl = bit_length(d, asz);
        if (bl == 0) {
            std::memset(o, 0, 4 * (size_t)L);
            exps[i] = ez;
            continue;
        }
        const int64_t s = bl > NB ? bl - NB : 0;           // right shift
        for (int64_t k = 0; k < rl; k++) r[k] = bits_from(d, asz, s + 64 * k);
        if (s > 0 && bit_at(d, asz, s - 1))                // half up on the magnitude
            for (int64_t k = 0; k < rl && ++r[k] == 0; k++) {
            }
        // L words of the magnitude, negated in two's complement for negative values
        uint32_t carry = 1;
        for (int k = 0; k < L; k++) {
            uint32_t w = (uint32_t)(r[k >> 1] >> (32 * (k & 1)));
            if (size < 0) {
                const uint64_t t = (uint64_t)(uint32_t)~w + carry;
                w = (uint32_t)t;
                carry = (uint32_t)(t >> 32);
            }
            o[k] = (int32_t)w;
        }
        exps[i] = 64 * (ex - asz) + s;
    }
    result = pos;
done:
    delete[] d;
    delete[] r;
    return result;
}

int64_t sdpv_encode_bigfloats(const int32_t* in, const int64_t* exps, int64_t n, int32_t L, int32_t nl,
                              uint8_t* out, int64_t cap, int64_t ez) {
    if (n < 0 || L < 1 || nl < 1 || !out || (n && (!in || !exps))) return -1;
    const int64_t ml = (L + 1) / 2 + 2;                    // 64-bit limbs: magnitude plus the shift
    uint64_t* m = new uint64_t[ml];
    int64_t pos = 0, result = -1;
    for (int64_t i = 0; i < n; i++) {
        const int32_t* x = in + i * (int64_t)L;
        const bool neg = x[L - 1] < 0;
        std::memset(m, 0, 8 * (size_t)ml);
        uint32_t carry = 1;
        for (int k = 0; k < L; k++) {                      // magnitude
            uint32_t w = (uint32_t)x[k];
            if (neg) {
                const uint64_t t = (uint64_t)(uint32_t)~w + carry;
                w = (uint32_t)t;
                carry = (uint32_t)(t >> 32);
            }
            m[k >> 1] |= (uint64_t)w << (32 * (k & 1));
        }
        if (exps[i] == ez || bit_length(m, ml) == 0) {
            if (pos + 1 > cap) goto done;
            out[pos++] = 1;
            continue;
        }
        // value = M 2^e = (M << r) 2^(64 q) with r = e mod 64 in [0, 64), q = (e - r) / 64
        const int64_t e = exps[i];
        const int64_t r = ((e % 64) + 64) % 64;
        const int64_t q = (e - r) / 64;
        if (r) {
            for (int64_t k = ml - 1; k > 0; k--) m[k] = (m[k] << r) | (m[k - 1] >> (64 - r));
            m[0] <<= r;
        }
        const int64_t size = (bit_length(m, ml) + 63) / 64;
        if (size > nl) goto done;
        if (pos + 13 + 8 * (int64_t)nl > cap) goto done;
        const int32_t ssize = (int32_t)(neg ? -size : size);
        const int64_t X = q + size;
        out[pos++] = 0;
        std::memcpy(out + pos, &ssize, 4);
        std::memcpy(out + pos + 4, &X, 8);
        pos += 12;
        std::memset(out + pos, 0, 8 * (size_t)nl);
        std::memcpy(out + pos, m, 8 * (size_t)size);
        pos += 8 * (int64_t)nl;
    }
    result = pos;
done:
    delete[] m;
    return result;
}

}  // extern "C"
