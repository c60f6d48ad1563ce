// Device helpers shared by the resolve passes: surface decode from a
// vis-buffer key and analytic barycentrics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
#include "internal.h"

namespace glimp {

// Surface of a covered key: meshlet (MeshletOffset + MeshletId of the draw
// that wrote it) and primitive id (Shading.cpp:328), from the rank in the
// key's low word and the framebuffer's draw table.
__device__ __forceinline__ void decode_surface(unsigned long long key, const DrawEntry* draws, uint32_t num_draws,
                                               uint32_t& meshlet, uint32_t& tri) {
    const uint32_t rank = ~(uint32_t)key;
    const uint32_t seq = rank >> 11;
    tri = ((rank >> 8) & 7u) * 16u + ((rank >> 3) & 15u);
    uint32_t lo = 0, hi = num_draws;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (draws[mid].seq_base <= seq) lo = mid; else hi = mid;
    }
    meshlet = draws[lo].meshlet_offset + (seq - draws[lo].seq_base);
}

// IntersectTriangle in the operation order of the reference's fast-math
// build (see the oracle for the derivation from the compiled shape). With
// Exact, the reciprocals are correctly rounded IEEE divisions instead of the
// reference's rcp14 + Newton step (the differentiable resolve uses this).
template <bool Exact = false>
__device__ inline void intersect_triangle(const ApproxTables& T, const float c[3][4], float su, float sv, float kx,
                                          float ky, float bary[3], float ddx[3], float ddy[3]) {
    auto frcp = [&](float x) { return Exact ? 1.0f / x : glimp::frcp(T, x); };
    float iw0 = frcp(c[0][3]), iw1 = frcp(c[1][3]), iw2 = frcp(c[2][3]);
    float p0x = c[0][0] * iw0, p0y = c[0][1] * iw0;
    float p1x = c[1][0] * iw1, p1y = c[1][1] * iw1;
    float p2x = c[2][0] * iw2, p2y = c[2][1] * iw2;
    float m0x = p2x - p1x, m1x = p0x - p1x, m1y = p0y - p1y;
    float d12y = p1y - p2y;
    float det = __builtin_fmaf(m0x, m1y, d12y * m1x);
    float invDet = frcp(det);
    float a0 = invDet * iw0, a1 = invDet * iw1, a2 = invDet * iw2;
    float dx0 = a0 * d12y, dx1 = a1 * (p2y - p0y), dx2 = a2 * m1y;
    float dy0 = a0 * m0x, dy1 = a1 * (p0x - p2x), dy2 = a2 * (p1x - p0x);
    float dxs = (dx0 + dx1) + dx2, dys = (dy0 + dy1) + dy2;
    float r0x = su - p0x, r0y = sv - p0y;
    float iiw = __builtin_fmaf(dys, r0y, __builtin_fmaf(dxs, r0x, iw0));
    float iW = frcp(iiw);
    float by = iW * __builtin_fmaf(dx1, r0x, dy1 * r0y);
    float bz = iW * __builtin_fmaf(dx2, r0x, dy2 * r0y);
    float bx = 1.0f - (by + bz);
    float wdx = frcp(__builtin_fmaf(kx, dxs, iiw)), wdy = frcp(__builtin_fmaf(ky, dys, iiw));
    float b[3] = { bx, by, bz }, dxk[3] = { dx0, dx1, dx2 }, dyk[3] = { dy0, dy1, dy2 };
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float bi = b[k] * iiw;
        ddx[k] = __builtin_fmaf(wdx, __builtin_fmaf(kx, dxk[k], bi), -b[k]);
        ddy[k] = __builtin_fmaf(wdy, __builtin_fmaf(ky, dyk[k], bi), -b[k]);
        bary[k] = b[k];
    }
}

}  // namespace glimp
