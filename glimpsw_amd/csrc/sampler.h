// Texture sampling shared by the resolve and the alpha test:
// Texture2D<RGBA8u>::SampleLevel with the SurfaceSampler (Repeat wrap,
// linear magnification, nearest minification; Shading.cpp:6-10).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"
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

// One channel of SampleLevel's bilinear filter. The reference lerps bytes
// held in 16-bit lanes, a + mulhrs(b - a, f << 7) (lerp16 above); with
// |b - a| <= 255 and f < 256 that is a + floor((b - a) * f / 256 + 1/2)
// (mulhrs rounds half up), and every term of it is exact in fp32: one
// subtract, one FMA, a floor and an add per channel instead of extracting,
// multiplying and repacking 16-bit halves.
__device__ __forceinline__ float lerp_byte(float a, float b, float fs) {
    return a + floorf(__builtin_fmaf(b - a, fs, 0.5f));
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
    const uint32_t d00 = D[i00], d10 = D[i00 + 8], d01 = D[i01], d11 = D[i01 + 8];
    const float fxs = inX ? (float)(xf & 255) * (1.0f / 256) : 0.0f;
    const float fys = (float)(yf & 255) * (1.0f / 256);
    uint32_t out = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; c++) {
        const uint32_t sh = 8 * c;
        const float r1 = lerp_byte((float)((d00 >> sh) & 255u), (float)((d10 >> sh) & 255u), fxs);
        const float r2 = lerp_byte((float)((d01 >> sh) & 255u), (float)((d11 >> sh) & 255u), fxs);
        out = __builtin_amdgcn_cvt_pk_u8_f32(lerp_byte(r1, r2, fys), c, out);
    }
    return out;
}

// SkyboxTex->SampleOctLevel<EnvSampler>(dir, 1) (Shading.cpp:677,
// Texture.h:467-480, EnvSampler Shading.cpp:11-15): MapOctahedron of the
// view direction (1/(|x|+|y|+|z|) from VRCP14PS alone, max(-z * w, 0) zero
// unless positive, c * 0.5 + 0.5 one FMA), ClampToEdge, level 1 (the mip
// fraction is 0), the float bilinear filter of SampleLinear on
// R11G11B10f texels (Texture.h:557-571, 157-182) with rowA + (rowB - rowA)
// * fy evaluated as fma(fy, fma(fx, c11 - c01, c01 - rowA), rowA).
__device__ inline void sample_sky(const ApproxTables& T, const DevTexture& tex, const float d[3], float out[3]) {
    const float w = rcp14(T, (fabsf(d[0]) + fabsf(d[1])) + fabsf(d[2]));
    const float mz = -d[2] * w;
    const float t = mz > 0.0f ? mz : 0.0f;
    const float sx = bitsf(fbits(t) ^ (fbits(d[0]) & 0x80000000u));
    const float sy = bitsf(fbits(t) ^ (fbits(d[1]) & 0x80000000u));
    const float u = __builtin_fmaf(__builtin_fmaf(d[0], w, sx), 0.5f, 0.5f);
    const float v = __builtin_fmaf(__builtin_fmaf(d[1], w, sy), 0.5f, 0.5f);
    int32_t ix = min(max(cvt_rne(u * tex.scale_lerp_u), 0), tex.mask_lerp_u);
    int32_t iy = min(max(cvt_rne(v * tex.scale_lerp_v), 0), tex.mask_lerp_v);
    const int32_t m = min(1, (int32_t)tex.mip_levels - 1);
    const int32_t stride = (int32_t)tex.row_shift - m;
    const int32_t offset = (int32_t)tex.mip_offsets[m];
    ix >>= m;
    iy >>= m;
    const int32_t xf = max(ix - 127, 0), yf = max(iy - 127, 0);
    const int32_t tx = xf >> 8, ty = yf >> 8;
    const bool inX = ((tx + 1) << m) < (int32_t)tex.width;
    const bool inY = ((ty + 1) << m) < (int32_t)tex.height;
    const int32_t i00 = offset + (int32_t)texel_offset(tx, ty, stride);
    const int32_t i01 = offset + (int32_t)texel_offset(tx, ty + (inY ? 1 : 0), stride);
    const uint32_t q[4] = { tex.data[i00], tex.data[i00 + 8], tex.data[i01], tex.data[i01 + 8] };
    const float fx = inX ? (float)(xf & 255) * (1.0f / 256) : 0.0f;
    const float fy = (float)(yf & 255) * (1.0f / 256);
    float c[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k][0] = bitsf((((q[k] >> 21) << 17) & 0x0FFE0000u) + 0x38000000u);
        c[k][1] = bitsf((((q[k] >> 10) << 17) & 0x0FFE0000u) + 0x38000000u);
        c[k][2] = bitsf(((q[k] << 18) & 0x0FFC0000u) + 0x38000000u);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        const float rowA = __builtin_fmaf(fx, c[1][ch] - c[0][ch], c[0][ch]);
        out[ch] = __builtin_fmaf(fy, __builtin_fmaf(fx, c[3][ch] - c[2][ch], c[2][ch] - rowA), rowA);
    }
}

}  // namespace glimp
