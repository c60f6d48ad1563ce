// Device-side arithmetic that reproduces the reference renderer's x86
// results bit for bit on gfx950.
//
// The reference is compiled with -ffast-math for AVX-512
// (src/SwRast/CMakeLists.txt:6). Its results therefore depend on
//  - VRCP14PS / VRSQRT14PS, called explicitly (SIMD.h:294-295) and emitted
//    for every vector float division (with one Newton-Raphson step);
//  - CVTPS2DQ rounding and its out-of-range result (SIMD.h:304);
//  - FMA contraction choices of the compiler.
// The two approximation instructions are table driven: a 16-bit mantissa
// looked up by the top 16 (rcp) or 15 (rsqrt) bits of the input mantissa
// reproduces them over all 2^32 inputs. Those tables are exactly piecewise
// linear over runs of 1024 entries, so the device evaluates them from 64 + 64
// integer pairs (A, B): entry j of run i is (A_i - B_i * j) >> K. The 1 KiB
// of pairs stays in the L1 cache.
//
// Device code is compiled with -ffp-contract=off: every FMA here is written
// out with __builtin_fmaf, every other operation rounds on its own.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace glimp {

struct ApproxTables {
    // [128] (A, -B) runs: 64 rcp runs, then 32 + 32 rsqrt runs (exponent parity 0, 1)
    const uint2* pwl;
};

// Entry j of a run: (A - B * j) >> shift, with -B stored. |B| < 2^17 and
// j < 2^10: a full-rate 24-bit multiply-add is exact (and wraps like the
// 32-bit table arithmetic).
__device__ __forceinline__ uint32_t pwl_entry(uint2 ab, uint32_t j, uint32_t shift) {
    return ((uint32_t)__mul24((int32_t)ab.y, (int32_t)j) + ab.x) >> shift;
}

// 16-bit result mantissa of VRCP14PS for the top 16 input mantissa bits.
__device__ __forceinline__ uint32_t rcp14_mant(const ApproxTables& T, uint32_t idx16) {
    return pwl_entry(T.pwl[idx16 >> 10], idx16 & 1023u, 12);
}

// 16-bit result mantissa of VRSQRT14PS for exponent parity `odd` and the
// top 15 input mantissa bits.
__device__ __forceinline__ uint32_t rsqrt14_mant(const ApproxTables& T, uint32_t odd, uint32_t idx15) {
    return pwl_entry(T.pwl[64 + odd * 32 + (idx15 >> 10)], idx15 & 1023u, 16);
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// Full VRCP14PS semantics (zeros, denormals, infinities, NaN, results that
// overflow or underflow); rcp14 only calls it off the fast path.
__device__ __forceinline__ float rcp14_special(const ApproxTables& T, float xf) {
    uint32_t x = fbits(xf);
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) return bitsf(m ? (x | 0x400000u) : s);
    if (e == 0) {
        if (m == 0) return bitsf(s | 0x7F800000u);
        int lz = __clz(m) - 8;          // shift so bit 23 is set
        m <<= lz;
        e = 1 - lz;
        m &= 0x7FFFFFu;
    }
    int re;
    uint32_t mant;
    if (m == 0) {
        re = 254 - e;
        mant = 0;
    } else {
        re = 253 - e;
        mant = rcp14_mant(T, m >> 7) << 7;
    }
    if (re >= 255) return bitsf(s | 0x7F800000u);
    if (re <= 0) {
        int sh = 1 - re;
        if (sh > 24) return bitsf(s);
        return bitsf(s | ((mant | 0x800000u) >> sh));
    }
    return bitsf(s | ((uint32_t)re << 23) | mant);
}

// VRCP14PS. Inputs with a biased exponent e in [1, 252] have a normal
// result: exponent 253 - e and the table mantissa, or 254 - e and mantissa 0
// for an exact power of two (a mantissa of 2^16 carries into the
// exponent); the bits are (253 << 23) - (x's exponent bits) + (mantissa <<
// 7), then x's sign. Everything else takes the general path.
__device__ __forceinline__ float rcp14(const ApproxTables& T, float xf) {
    const uint32_t x = fbits(xf);
    // d = x's exponent bits - (253 << 23), mod 2^32: exponents 1-252 (the
    // fast path) give d in [0x82000000, 0xFF800000], and the result's bits
    // are (mantissa << 7) - d
    const uint32_t d = (x & 0x7F800000u) + 0x81800000u;
    if (__builtin_expect(d <= 0x81FFFFFFu, 0)) return rcp14_special(T, xf);
    uint32_t mant = rcp14_mant(T, (x >> 7) & 0xFFFFu);
    mant = (x & 0x7FFFFFu) ? mant : 0x10000u;
    const uint32_t r = (mant << 7) - d;
    return bitsf((r & 0x7FFFFFFFu) | (x & 0x80000000u));
}

// Full VRSQRT14PS semantics; rsqrt14 only calls it off the fast path.
__device__ __forceinline__ float rsqrt14_special(const ApproxTables& T, float xf) {
    uint32_t x = fbits(xf);
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) {
        if (m) return bitsf(x | 0x400000u);
        return bitsf(s ? 0xFFC00000u : 0u);
    }
    if (e == 0) {
        if (m == 0) return bitsf(s | 0x7F800000u);
        int lz = __clz(m) - 8;
        m <<= lz;
        e = 1 - lz;
        m &= 0x7FFFFFu;
    }
    if (s) return bitsf(0xFFC00000u);
    int ue = e - 127;
    int oc = ue & 1;
    int half = (ue - oc) / 2;
    if (m == 0 && oc == 0) return bitsf((uint32_t)(127 - half) << 23);
    return bitsf(((uint32_t)(126 - half) << 23) | (rsqrt14_mant(T, (uint32_t)oc, m >> 8) << 7));
}

// VRSQRT14PS. Positive normal inputs take the fast path: for x = 2^u * 1.m
// the result exponent is 126 - floor(u / 2) = 190 - (e + 1) / 2 (e the biased
// exponent), the table is chosen by u's parity (bit 23 of x, inverted) and the
// top mantissa bits, and an even power of two (m = 0, u even) gives mantissa
// 2^16, which carries into the exponent (127 - u / 2).
__device__ __forceinline__ float rsqrt14(const ApproxTables& T, float xf) {
    const uint32_t x = fbits(xf);
    if (__builtin_expect(x - 0x00800000u >= (254u << 23), 0)) return rsqrt14_special(T, xf);
    uint32_t mant = pwl_entry(T.pwl[64 + (((x >> 18) & 63u) ^ 32u)], (x >> 8) & 1023u, 16);
    mant = (x & 0x00FFFFFFu) == 0x00800000u ? 0x10000u : mant;
    return bitsf((0x5F000000u - (((x + 0x00800000u) >> 1) & 0x7F800000u)) + (mant << 7));
}

// Vector a/b as the reference's fast-math build computes it:
// r = rcp14(b); q = a*r; e = fma(q, b, -a); q - r*e (fused).
__device__ __forceinline__ float fdiv(const ApproxTables& T, float a, float b) {
    float r = rcp14(T, b);
    float q = a * r;
    float e = __builtin_fmaf(q, b, -a);
    return __builtin_fmaf(-r, e, q);
}
__device__ __forceinline__ float frcp(const ApproxTables& T, float b) { return fdiv(T, 1.0f, b); }

// CVTPS2DQ: round to nearest even; NaN and out-of-range give INT32_MIN.
__device__ __forceinline__ int32_t cvt_rne(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)__builtin_rintf(x);
}

// cvttps2dq: truncation toward zero; out of range and NaN give INT32_MIN.
__device__ __forceinline__ int32_t cvt_trunc(float x) {
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}

// Packed signed 16-bit pairs: x in the low half, y in the high half.
__device__ __forceinline__ int32_t lo16(uint32_t v) { return (int32_t)(int16_t)(v & 0xFFFFu); }
__device__ __forceinline__ int32_t hi16(uint32_t v) { return (int32_t)(int16_t)(v >> 16); }
__device__ __forceinline__ uint32_t pk16(int32_t lo, int32_t hi) {
    return (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16);
}

// simd::mul(mat4, float4(p, 1)) component c (SIMD.h:457-464), m column-major.
__device__ __forceinline__ float xform(const float* m, int c, float x, float y, float z) {
    return __builtin_fmaf(x, m[c], __builtin_fmaf(y, m[4 + c], __builtin_fmaf(z, m[8 + c], m[12 + c])));
}

// IEEE binary16 -> binary32, exact (what VCVTPH2PS computes).
// n / d and n % d for n < 2^24, 1 <= d < 2^24: a float estimate within one
// of the quotient, then one correction step (cheaper than an integer divide).
__device__ __forceinline__ uint32_t div_small(uint32_t n, uint32_t d, uint32_t& rem) {
    uint32_t q = (uint32_t)((float)n * __builtin_amdgcn_rcpf((float)d));
    int32_t r = (int32_t)n - (int32_t)__umul24(q, d);
    if (r < 0) { q--; r += (int32_t)d; }
    else if (r >= (int32_t)d) { q++; r -= (int32_t)d; }
    rem = (uint32_t)r;
    return q;
}

// Half to float (exact, like VCVTPH2PS).
__device__ __forceinline__ float half_to_float(uint32_t h) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}

}  // namespace glimp
