// Device helpers for surface normals shared by the resolve and the deferred
// pipeline's G-buffer encode: vector helpers as the reference's SIMD layer
// defines them (SIMD.h:427-439), the vertex normal/tangent decode, the
// interpolation into world space and the normal mapping, in the operation
// order of the reference's optimised build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"

namespace glimp {

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return __builtin_fmaf(a[0], b[0], __builtin_fmaf(a[1], b[1], a[2] * b[2]));
}

__device__ __forceinline__ void normalize3(const ApproxTables& T, float v[3]) {
    float r = rsqrt14(T, dot3(v, v));
    v[0] = v[0] * r;
    v[1] = v[1] * r;
    v[2] = v[2] * r;
}

__device__ __forceinline__ float mulsign(float x, float y) {
    return bitsf(fbits(x) ^ (fbits(y) & 0x80000000u));
}

__device__ __forceinline__ float lerpf(float a, float b, float t) {
    return __builtin_fmaf(t, b, __builtin_fmaf(-t, a, a));
}

// normalize(mul(ObjectToWorldMat, BaryLerp(bary, a0, a1, a2)))
// (Shading.cpp:546-554, BaryLerp Rasterizer.h:101-104)
__device__ __forceinline__ void lerp_to_world(const ApproxTables& T, const float a[3][3], const float bary[3],
                                              const float* M, float out[3]) {
    float ai[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
        ai[j] = __builtin_fmaf(a[0][j], bary[0], __builtin_fmaf(a[1][j], bary[1], a[2][j] * bary[2]));
#pragma unroll
    for (int j = 0; j < 3; j++)
        out[j] = __builtin_fmaf(ai[0], M[j], __builtin_fmaf(ai[1], M[3 + j], ai[2] * M[6 + j]));
    normalize3(T, out);
}

// Normal mapping of ResolveSurface (Shading.cpp:556-572): bitangent =
// cross(N, T) with the handedness sign, tangent-space normal from the NMR
// texel's two 8-bit channels with z = approx_sqrt(1 - x^2 - y^2), result
// normalize(xT + yB + zN), in the operation order of the reference's
// optimised build (`c * (1/127.5) - 1` and the sums contracted into FMAs,
// 1 - (x^2 + y^2) evaluated as 1 + fma(-x, x, -y^2)).
__device__ __forceinline__ void normal_map(const ApproxTables& T, float n[3], const float t[3], uint32_t handed,
                                           uint32_t nmr) {
    float b[3] = { __builtin_fmaf(n[1], t[2], -(n[2] * t[1])), __builtin_fmaf(n[2], t[0], -(n[0] * t[2])),
                   __builtin_fmaf(n[0], t[1], -(n[1] * t[0])) };
#pragma unroll
    for (int k = 0; k < 3; k++) b[k] = bitsf(fbits(b[k]) ^ handed);
    const float sx = __builtin_fmaf((float)(nmr & 255u), 1.0f / 127.5f, -1.0f);
    const float sy = __builtin_fmaf((float)((nmr >> 8) & 255u), 1.0f / 127.5f, -1.0f);
    const float s = 1.0f + __builtin_fmaf(-sx, sx, -(sy * sy));
    const float sz = rsqrt14(T, s) * s;
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] = __builtin_fmaf(sz, n[k], __builtin_fmaf(sy, b[k], sx * t[k]));
    const float inv = rsqrt14(T, __builtin_fmaf(r[0], r[0], __builtin_fmaf(r[1], r[1], r[2] * r[2])));
#pragma unroll
    for (int k = 0; k < 3; k++) n[k] = r[k] * inv;
}


// UnpackNormalTangent (Shading.cpp:232-236): normal from bytes 0-1, tangent
// (shift 16) from bytes 2-3.
__device__ __forceinline__ void unpack_normal(const ApproxTables& T, uint32_t packed, float n[3]) {
    const float scale = 1.0f / 255;
    float u = (float)(packed & 255) * scale, v = (float)((packed >> 8) & 255) * scale;
    u = u * 2.0f - 1.0f;
    v = v * 2.0f - 1.0f;
    n[0] = u;
    n[1] = v;
    n[2] = (1.0f - fabsf(u)) - fabsf(v);
    float t = fmaxf(-n[2], 0.0f);
    n[0] = n[0] - mulsign(t, n[0]);
    n[1] = n[1] - mulsign(t, n[1]);
    normalize3(T, n);
}

}  // namespace glimp
