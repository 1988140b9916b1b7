// pt_libm.h -- the Fresnel term's power x^5 as the reference's host build
// computes it, for the device and the host.
//
// The reference evaluates Schlick's approximation with pow(1-cosine, 5)
// (src/renderer/transport.h:147); its host build promotes pow(float, int) to
// pow(double, double), i.e. the host libm's pow.  That pow is not correctly
// rounded everywhere (glibc 2.35 differs from the correctly rounded value by
// one unit in the last place on about 1 float input in 2000), so the device
// evaluates x^5 correctly rounded, exactly (pow5_rounded), and the runtime
// replaces the inputs where the host libm differs by the host's values
// (pow5_table: built once per process by the host over every x = 1 - c the
// shading can form with c in [0, 4], i.e. glass up to refractive index 4).
#pragma once

#include <stdint.h>

#ifndef PT_HD
#define PT_HD __host__ __device__ __forceinline__
#endif

namespace pt {

// The double nearest to x^5 (ties to even), for finite float x.  x = m * 2^e
// with an integer m < 2^24, so m^5 < 2^120 is formed exactly in two 64-bit
// words and rounded to 53 bits; x^5 of any float is a normal double.
PT_HD double pow5_rounded(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    const uint64_t sign = uint64_t(b >> 31) << 63;
    int e = int((b >> 23) & 255u);
    uint64_t m = b & 0x7fffffu;
    if(e == 0) {
        if(m == 0) return __builtin_bit_cast(double, sign);
        e = 1;
    } else {
        m |= 0x800000u;
    }
    // x = m * 2^(e - 150); m^5 = (hi:lo)
    const uint64_t m2 = m * m;                                  // < 2^48
    const unsigned __int128 m5 = (unsigned __int128)m2 * m2 * m;  // < 2^120
    const uint64_t hi = uint64_t(m5 >> 64), lo = uint64_t(m5);
    int exp2 = 5 * (e - 150);
    // bit length of m^5
    const int bits = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(lo);
    uint64_t mant;
    if(bits > 53) {
        const int shift = bits - 53;
        uint64_t q = uint64_t(m5 >> shift);
        const unsigned __int128 rem = m5 - ((unsigned __int128)q << shift);
        const unsigned __int128 half = (unsigned __int128)1 << (shift - 1);
        if(rem > half || (rem == half && (q & 1u))) q++;
        exp2 += shift;
        if(q >> 53) { q >>= 1; exp2++; }
        mant = q;
    } else {
        mant = lo << (53 - bits);
        exp2 -= 53 - bits;
    }
    // value = mant * 2^exp2 with 2^52 <= mant < 2^53
    const uint64_t biased = uint64_t(exp2 + 52 + 1023);
    return __builtin_bit_cast(double, sign | (biased << 52) | (mant & ((uint64_t(1) << 52) - 1)));
}

// Table domain: x = k * 2^-24 for integer k in [-3 * 2^24, 2^24], which holds
// every float 1 - c for a float c in [0, 4].
constexpr int32_t pow5_table_low = -3 * (1 << 24);
constexpr int32_t pow5_table_high = 1 << 24;

// Index k of x in the table domain, or false when x is outside it.
PT_HD bool pow5_table_index(float x, int32_t& k) {
    if(!(x >= -3.0f && x <= 1.0f)) return false;
    const float scaled = x * 0x1p24f;  // exact
    const int32_t i = int32_t(scaled);
    if(float(i) != scaled) return false;
    k = i;
    return true;
}

// Host values where they differ from pow5_rounded: keys (k, ascending) and values.
struct pow5_table {
    const int32_t *keys;
    const double *values;
    int32_t count;
};

// x^5 as the host libm's pow(double(x), 5.0) returns it.
PT_HD double libm_pow5(const pow5_table& t, float x) {
    int32_t k;
    if(t.count > 0 && pow5_table_index(x, k)) {
        int32_t lo = 0, hi = t.count;
        while(lo < hi) {
            const int32_t mid = lo + (hi - lo) / 2;
            if(t.keys[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        if(lo < t.count && t.keys[lo] == k) return t.values[lo];
    }
    return pow5_rounded(x);
}

} // namespace pt
