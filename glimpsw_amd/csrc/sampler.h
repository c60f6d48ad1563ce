// Texture sampling shared by the resolve and the alpha test:
// Texture2D<RGBA8u>::SampleLevel with the SurfaceSampler (Repeat wrap,
// linear magnification, nearest minification; Shading.cpp:6-10).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace glimp {

__device__ __forceinline__ uint32_t texel_offset(int32_t x, int32_t y, int32_t stride) {  // TiledY8
    return (uint32_t)((y & 7) | (x << 3) | ((y & ~7) << stride));
}

// simd::lerp16 on the two 16-bit halves: a + mulhrs(b - a, t).
__device__ __forceinline__ uint32_t lerp16(uint32_t a, uint32_t b, uint32_t t) {
    uint32_t r = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        int32_t av = (int16_t)(a >> (16 * h)), bv = (int16_t)(b >> (16 * h)), tv = (int16_t)(t >> (16 * h));
        int32_t d = (int16_t)(bv - av);
        int32_t prod = __mul24(d, tv);   // 16-bit operands: exact at full rate
        int32_t mh = (int16_t)(((prod >> 14) + 1) >> 1);
        r |= (uint32_t)(uint16_t)(int16_t)(av + mh) << (16 * h);
    }
    return r;
}

// Texture2D<RGBA8u>::SampleLevel with Repeat wrap (Texture.h:412-459,
// 506-556); `linear` is the tile's filter choice.
__device__ inline uint32_t sample_level(const DevTexture& tex, float u, float v, int32_t mipIn, bool linear,
                                       uint32_t layer = 0) {
    float su = u * tex.scale_lerp_u, sv = v * tex.scale_lerp_v;
    int32_t ix = cvt_rne(su) & tex.mask_lerp_u;
    int32_t iy = cvt_rne(sv) & tex.mask_lerp_v;
    int32_t m = min(max(mipIn, 0), (int32_t)tex.mip_levels - 1);
    int32_t offset = (int32_t)(layer * tex.layer_stride + tex.mip_offsets[m]);
    int32_t stride = (int32_t)tex.row_shift - m;
    int32_t x = ix >> m, y = iy >> m;
    const uint32_t* D = tex.data;
    if (!linear) return D[offset + (int32_t)texel_offset(x >> 8, y >> 8, stride)];
    int32_t xf = max(x - 127, 0), yf = max(y - 127, 0);
    int32_t tx = xf >> 8, ty = yf >> 8;
    bool inX = ((tx + 1) << m) < (int32_t)tex.width;
    bool inY = ((ty + 1) << m) < (int32_t)tex.height;
    int32_t i00 = offset + (int32_t)texel_offset(tx, ty, stride);
    int32_t i01 = offset + (int32_t)texel_offset(tx, ty + (inY ? 1 : 0), stride);
    uint32_t d00 = D[i00], d10 = D[i00 + 8], d01 = D[i01], d11 = D[i01 + 8];
    uint32_t fx = (uint32_t)(xf & 255) << 7, fy = (uint32_t)(yf & 255) << 7;
    fx = (fx << 16) | fx;
    fy = (fy << 16) | fy;
    if (!inX) fx = 0;
    uint32_t rb1 = lerp16(d00 & 0x00FF00FFu, d10 & 0x00FF00FFu, fx);
    uint32_t ga1 = lerp16((d00 >> 8) & 0x00FF00FFu, (d10 >> 8) & 0x00FF00FFu, fx);
    uint32_t rb2 = lerp16(d01 & 0x00FF00FFu, d11 & 0x00FF00FFu, fx);
    uint32_t ga2 = lerp16((d01 >> 8) & 0x00FF00FFu, (d11 >> 8) & 0x00FF00FFu, fx);
    return lerp16(rb1, rb2, fy) | (lerp16(ga1, ga2, fy) << 8);
}

}  // namespace glimp
