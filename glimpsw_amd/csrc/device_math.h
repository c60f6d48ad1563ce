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
// reproduces them over all 2^32 inputs. The tables live in device memory
// (128 KiB each, hot entries stay in L2/L1).
//
// Device code is compiled with -ffp-contract=off: every FMA here is written
// out with __builtin_fmaf, every other operation rounds on its own.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace glimp {

struct ApproxTables {
    const uint16_t* rcp;      // [65536]
    const uint16_t* rsqrt;    // [2][32768]
};

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

__device__ __forceinline__ float rcp14(const ApproxTables& T, float xf) {
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
        mant = (uint32_t)T.rcp[m >> 7] << 7;
    }
    if (re >= 255) return bitsf(s | 0x7F800000u);
    if (re <= 0) {
        int sh = 1 - re;
        if (sh > 24) return bitsf(s);
        return bitsf(s | ((mant | 0x800000u) >> sh));
    }
    return bitsf(s | ((uint32_t)re << 23) | mant);
}

__device__ __forceinline__ float rsqrt14(const ApproxTables& T, float xf) {
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
    return bitsf(((uint32_t)(126 - half) << 23) | ((uint32_t)T.rsqrt[oc * 32768 + (m >> 8)] << 7));
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
__device__ __forceinline__ float half_to_float(uint32_t h) {
    uint32_t s = (h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 0) {
        float f = (float)m * (1.0f / 16777216.0f);
        return s ? -f : f;
    }
    if (e == 31) return bitsf(s | 0x7F800000u | (m << 13));
    return bitsf(s | ((e + 112u) << 23) | (m << 13));
}

}  // namespace glimp
