// Vis-buffer resolve for gfx950: one lane per pixel, 16 lanes per 4x4 tile.
//
// ShadingContext::Resolve / ResolveDebug (Shading.cpp:658-773): decode the
// surface of each covered pixel from its 64-bit key, re-transform the
// triangle's three vertices, recover perspective-correct barycentrics and
// their screen derivatives analytically (IntersectTriangle,
// Shading.cpp:417-464), interpolate UVs, pick the mip level from the UV
// gradients and sample the material texture (Texture.h:412-575); in lit mode
// also interpolate the vertex normal (normal-mapped with a second texture
// layer) and evaluate the lights (Shading.cpp:602-645), tonemap, pack and
// blend the light sprites. The other debug layers (normals, metallic /
// roughness, hashed meshlet and triangle ids) share the surface code.
//
// The reference shades a 4x4 tile per SIMD step; two decisions are taken
// per tile: the texture filter (minification anywhere in the tile selects the
// min filter, Texture.h:432) and the per-light early-outs. A 64-lane wave
// holds four tiles, so those decisions are ballots over 16-lane groups.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "resolve.h"
#include "normals.h"
#include "sampler.h"
#include "surface.h"

namespace glimp {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// AlphaBlendU8 (Shading.cpp:238-246): source over in 8.8 fixed point.
__device__ __forceinline__ uint32_t alpha_blend_u8(uint32_t bg, uint32_t fg) {
    uint32_t t = (fg >> 1) & 0x7F800000u;
    t |= (t >> 16) | 0x00400040u;
    const uint32_t rb = lerp16(bg & 0x00FF00FFu, fg & 0x00FF00FFu, t);
    const uint32_t ag = lerp16((bg >> 8) & 0x00FF00FFu, (fg >> 8) & 0x00FF00FFu, t);
    return rb | (ag << 8);
}

__device__ __forceinline__ uint32_t pack_rgba8(float r, float g, float b, float a) {  // Texture.h:55-67
    float c[4] = { r, g, b, a };
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int32_t i = cvt_rne(c[k] * 255.0f);
        i = min(max(i, -32768), 32767);
        i = min(max(i, 0), 255);
        out |= (uint32_t)i << (8 * k);
    }
    return out;
}

__device__ __forceinline__ void unpack_srgb(uint32_t packed, float c[3]) {  // Texture.h:37-54
    uint32_t rb1 = ((packed << 8) & 0xFF00FF00u) + 0x00FF00FFu;
    uint32_t ag1 = (packed & 0xFF00FF00u) + 0x00FF00FFu;
    uint32_t rlo = rb1 & 0xFFFFu, rhi = rb1 >> 16, glo = ag1 & 0xFFFFu;
    const float scale = 1.0f / 65535;
    c[0] = (float)((rlo * rlo) >> 16) * scale;
    c[1] = (float)((glo * glo) >> 16) * scale;
    c[2] = (float)((rhi * rhi) >> 16) * scale;
}

__device__ float light_attenuation(const ApproxTables& T, const ResolveLight& L, const float pos[3]) {
    if (L.type == 0) return 1.0f;
    float d[3] = { L.position[0] - pos[0], L.position[1] - pos[1], L.position[2] - pos[2] };
    float d2 = dot3(d, d);
    float factor = d2 * L.inv_radius_sq;
    float smooth = fmaxf(1.0f - factor * factor, 0.0f);
    float att = (smooth * smooth) * rcp14(T, fmaxf(d2, 1e-4f));
    if (L.type == 2) {
        float nd[3] = { d[0], d[1], d[2] };
        normalize3(T, nd);
        float negDir[3] = { -L.direction[0], -L.direction[1], -L.direction[2] };
        float cd = dot3(negDir, nd);
        float sa = fminf(fmaxf(cd * L.spot_scale + L.spot_offset, 0.0f), 1.0f);
        att *= sa * sa;
    }
    return att;
}

// World-space surface normal of ResolveSurface (Shading.cpp:546-572): the
// interpolated vertex normal, normal-mapped when a lane of the tile has a
// normal texel. Called by every lane of the wave (the tile decision is a
// ballot); `active` lanes get their normal.
__device__ __forceinline__ void surface_normal(const ResolveParams& p, bool active, uint32_t mi, uint32_t tri,
                                               const float bary[3], uint32_t nmr, unsigned long long groupMask,
                                               float N[3]) {
    const ApproxTables& T = p.tables;
    if (active) {
        const uint8_t* I = p.m.indices + (size_t)mi * 384;
        float nv[3][3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t vi = I[k * 128 + tri];
            unpack_normal(T, p.m.normal_tangents[(size_t)mi * 64 + vi], nv[k]);
        }
        lerp_to_world(T, nv, bary, p.object_to_world, N);
    }
    if (__ballot((nmr & 0xFFFFu) != 0) & groupMask) {
        if (active) {
            const uint8_t* I = p.m.indices + (size_t)mi * 384;
            float tv[3][3], Tw[3];
#pragma unroll
            for (int k = 0; k < 3; k++) unpack_normal(T, p.m.normal_tangents[(size_t)mi * 64 + I[k * 128 + tri]] >> 16, tv[k]);
            lerp_to_world(T, tv, bary, p.object_to_world, Tw);
            // handedness of the first vertex (Shading.cpp:501-503)
            const uint32_t handed = ((p.m.tangent_handedness[mi] >> (I[tri] & 63u)) & 1u) << 31;
            normal_map(T, N, Tw, handed, nmr);
        }
    }
}

// Mode: GLIMP_RESOLVE_* (0 BaseColor, 1 lit, 2 Normals, 3 MetallicRoughness,
// 4 MeshletId, 5 TriangleId); each output is its own instantiation.
template <int Mode>
__global__ void __launch_bounds__(256) k_resolve(ResolveParams p) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t W = (uint32_t)p.width, H = (uint32_t)p.height;
    const uint32_t tilesX = W / 4;
    const uint32_t tile = t >> 4, q = t & 15u;
    uint32_t tx;
    const uint32_t ty = div_small(tile, tilesX, tx);
    const bool inside = ty * 4 < H;
    const uint32_t x = tx * 4 + (q & 3u), y = ty * 4 + (q >> 2);
    const uint32_t lane = lane_id();
    const unsigned long long groupMask = 0xFFFFull << (lane & 48u);
    const ApproxTables& T = p.tables;

    unsigned long long key = inside ? p.keys[t] : p.clear_key;
    const float depth = bitsf((uint32_t)(key >> 32));
    const bool active = inside && !(depth <= 0.0f);

    // surface of the pixel
    float gradU[2] = { 0, 0 }, gradV[2] = { 0, 0 }, U = 0, V = 0, bary[3] = { 0, 0, 0 };
    uint32_t material = 0xFFFFFFFFu, mi = 0, tri = 0;
    if (active) decode_surface(key, p.draws, p.num_draws, mi, tri);
    if (Mode < 4 && active) {
        const uint8_t* I = p.m.indices + (size_t)mi * 384;
        const float* P = p.m.positions + (size_t)mi * 192;
        float c[3][4], uv[3][2];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t vi = I[k * 128 + tri];
            float px = P[vi], py = P[64 + vi], pz = P[128 + vi];
#pragma unroll
            for (int j = 0; j < 4; j++) c[k][j] = xform(p.object_to_clip, j, px, py, pz);
            uint32_t tc = p.m.texcoords[(size_t)mi * 64 + vi];
            uv[k][0] = half_to_float(tc & 0xFFFFu);
            uv[k][1] = half_to_float(tc >> 16);
        }
        material = p.m.material_id[mi];
        const float su = __builtin_fmaf((float)(int32_t)x, p.scale_x, p.center_x);
        const float sv = __builtin_fmaf((float)(int32_t)y, p.scale_y, p.center_y);
        float ddx[3], ddy[3];
        intersect_triangle(T, c, su, sv, p.kx, p.ky, bary, ddx, ddy);
        float t10x = uv[1][0] - uv[0][0], t10y = uv[1][1] - uv[0][1];
        float t20x = uv[2][0] - uv[0][0], t20y = uv[2][1] - uv[0][1];
        U = __builtin_fmaf(bary[2], t20x, __builtin_fmaf(bary[1], t10x, uv[0][0]));
        V = __builtin_fmaf(bary[2], t20y, __builtin_fmaf(bary[1], t10y, uv[0][1]));
        gradU[0] = __builtin_fmaf(t20x, ddx[2], t10x * ddx[1]);
        gradV[0] = __builtin_fmaf(t20y, ddx[2], t10y * ddx[1]);
        gradU[1] = __builtin_fmaf(t20x, ddy[2], t10x * ddy[1]);
        gradV[1] = __builtin_fmaf(t20y, ddy[2], t10y * ddy[1]);
    }

    // material waterfall per 4x4 tile (Shading.cpp:532-545): each pass
    // settles one material of every tile (its texture, the pixel's mip level
    // and the tile's filter vote); the texture is sampled once afterwards
    uint32_t albedo = 0;
    const DevTexture* myTex = nullptr;
    int32_t myMip = 0;
    bool myLinear = false;
    bool done = Mode >= 4 || !active || material == 0xFFFFFFFFu || material >= p.num_materials;
    for (;;) {
        const unsigned long long todo = __ballot(!done);
        if (!todo) break;
        const unsigned long long mine = todo & groupMask;
        const uint32_t leader = mine ? (uint32_t)__ffsll((long long)mine) - 1 : lane;
        const uint32_t lmat = __shfl(material, (int)leader);
        int32_t texIdx = -1;
        if (mine) texIdx = p.materials[lmat].texture;
        int32_t mip = -64;
        // a reference into the table, not a local copy: a copy indexed by the
        // mip level would live in scratch memory
        const DevTexture* tex = texIdx >= 0 ? &p.textures[texIdx] : nullptr;
        if (tex) {
            float ssu = tex->scale_u * tex->scale_u, ssv = tex->scale_v * tex->scale_v;
            float dx = __builtin_fmaf(gradU[0], gradU[0], gradV[0] * gradV[0]) * ssu;
            float dy = __builtin_fmaf(gradU[1], gradU[1], gradV[1] * gradV[1]) * ssv;
            float mx = (dx > dy) ? dx : dy;
            mip = (int32_t)(fbits(mx) - (127u << 23)) >> 24;
        }
        const unsigned long long vote = __ballot(active && mip > 0);
        const bool minify = (vote & groupMask) != 0;
        if (!done && material == lmat) {
            myTex = tex;
            myMip = mip;
            myLinear = !minify;
            done = true;
        }
    }
    // SurfaceSampler: Repeat, Mag Linear, Min Nearest (Shading.cpp:6-10);
    // layer 1 holds normal xy, metallic and roughness (Shading.cpp:541-543)
    uint32_t nmr = 0;
    if (myTex) {
        albedo = sample_level(*myTex, U, V, myMip, myLinear);
        if (Mode >= 1 && Mode <= 3 && myTex->num_layers >= 2) nmr = sample_level(*myTex, U, V, myMip, myLinear, 1);
    }

    uint32_t outc;
    if constexpr (Mode != 1) {
        // ResolveDebug (Shading.cpp:734-773): the layer's colour, the
        // background checker where nothing was drawn
        const uint32_t x0 = tx * 4, y0 = ty * 4;
        const uint32_t bg = ((x0 ^ y0) & 4u) ? 0xFFA0A0A0u : 0xFFFFFFFFu;
        const float s = 1.0f / 255;
        float c[3] = { 0, 0, 0 };
        if constexpr (Mode == 0) {
            c[0] = (float)(albedo & 255) * s;
            c[1] = (float)((albedo >> 8) & 255) * s;
            c[2] = (float)((albedo >> 16) & 255) * s;
        } else if constexpr (Mode == 2) {
            float N[3] = { 0, 0, 0 };
            surface_normal(p, active, mi, tri, bary, nmr, groupMask, N);
#pragma unroll
            for (int k = 0; k < 3; k++) c[k] = __builtin_fmaf(N[k], 0.5f, 0.5f);
        } else if constexpr (Mode == 3) {
            c[0] = (float)((nmr >> 16) & 255u) * s;
            c[1] = (float)((nmr >> 24) & 255u) * s;
        } else {
            // RGBA8u::Unpack of the hashed surface id (MeshletId: id / 128)
            const uint32_t sid = mi * 128u + tri;
            const uint32_t h = (Mode == 4 ? sid / 128u : sid) * 123456789u;
            c[0] = (float)(h & 255) * s;
            c[1] = (float)((h >> 8) & 255) * s;
            c[2] = (float)((h >> 16) & 255) * s;
        }
        outc = active ? pack_rgba8(c[0], c[1], c[2], 1.0f) : bg;
    } else {
        float color[3] = { 0, 0, 0 };
        float N[3] = { 0, 0, 0 }, Vd[3] = { 0, 0, 0 }, wpos[3] = { 0, 0, 0 }, base[3] = { 0, 0, 0 };
        float f0[3] = { 0, 0, 0 }, diffuse[3] = { 0, 0, 0 }, NoV = 0, aR = 1e-4f;
        surface_normal(p, active, mi, tri, bary, nmr, groupMask, N);
        if (active) {
            float c4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) c4[j] = xform(p.inv_screen_proj, j, (float)(int32_t)x, (float)(int32_t)y, depth);
            float rw = frcp(T, c4[3]);
            wpos[0] = c4[0] * rw; wpos[1] = c4[1] * rw; wpos[2] = c4[2] * rw;
            unpack_srgb(albedo, base);
            // SurfaceData::Metallic / Roughness (Shading.cpp:574-577), 0 without an NMR layer
            const float metal = (float)((nmr >> 16) & 255u) * (1.0f / 255);
            const float rough = (float)((nmr >> 24) & 255u) * (1.0f / 255);
            aR = fmaxf(rough * rough, 1e-4f);
            const float refl = 0.16f * 0.5f * 0.5f;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                f0[k] = lerpf(refl, base[k], metal);
                diffuse[k] = base[k] * (1.0f - metal);
                Vd[k] = p.view_pos[k] - wpos[k];
            }
            normalize3(T, Vd);
            NoV = fabsf(dot3(N, Vd)) + 1e-5f;
        }
        for (uint32_t li = 0; li < p.num_lights; li++) {
            const ResolveLight& L = p.lights[li];
            float Ld[3] = { 0, 0, 0 }, NoL = 0, att = 0;
            if (active) {
                if (L.type == 0) {
                    Ld[0] = -L.direction[0]; Ld[1] = -L.direction[1]; Ld[2] = -L.direction[2];
                } else {
                    Ld[0] = L.position[0] - wpos[0]; Ld[1] = L.position[1] - wpos[1]; Ld[2] = L.position[2] - wpos[2];
                    normalize3(T, Ld);
                }
                NoL = dot3(N, Ld);
            }
            if (!(__ballot(active && !(NoL < 1e-4f)) & groupMask)) continue;
            if (active) att = light_attenuation(T, L, wpos) * L.intensity * p.light_exposure;
            if (!(__ballot(active && !(NoL * att < 1e-4f)) & groupMask)) continue;
            if (active) {
                float Hv[3] = { Vd[0] + Ld[0], Vd[1] + Ld[1], Vd[2] + Ld[2] };
                normalize3(T, Hv);
                float NoH = fminf(fmaxf(dot3(N, Hv), 0.0f), 1.0f);
                float LoH = fminf(fmaxf(dot3(Ld, Hv), 0.0f), 1.0f);
                float a = NoH * aR;
                float kk = aR * rcp14(T, (1.0f - NoH * NoH) + a * a);
                float D = kk * kk * 0.3183098861837907f;
                float va = 2.0f * NoL * NoV;
                float vb = NoL + NoV;
                float Vis = fdiv(T, 0.5f, lerpf(va, vb, aR));
                float u1 = 1.0f - LoH;
                float f = (u1 * u1) * (u1 * u1) * u1;
                float w = fmaxf(NoL * att, 0.0f);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float F = f + f0[c] * (1.0f - f);
                    float Fr = (D * Vis) * F;
                    float Fd = diffuse[c] * 0.3183098861837907f;
                    color[c] = color[c] + (Fd + Fr) * L.color[c] * w;
                }
            }
        }
        float cc[3];
        // sky: the view direction through the pixel at depth 1
        // (Shading.cpp:663-667, 675-678)
        float sky[3] = { 0, 0, 0 };
        if (!active && inside && p.has_sky) {
            float c4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) c4[j] = xform(p.inv_screen_proj, j, (float)(int32_t)x, (float)(int32_t)y, 1.0f);
            const float rw = frcp(T, c4[3]);
            const float dir[3] = { c4[0] * rw - p.view_pos[0], c4[1] * rw - p.view_pos[1], c4[2] * rw - p.view_pos[2] };
            sample_sky(T, p.sky, dir, sky);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float v = active ? (color[k] + base[k] * 0.05f) * p.exposure : sky[k] * p.exposure;
            cc[k] = fdiv(T, v, v + 0.155f) * 1.019f;
        }
        outc = pack_rgba8(cc[0], cc[1], cc[2], 1.0f);
        // Light sprites (Shading.cpp:689-727), blended in light order over the
        // pixel's colour: the reference's pass after the image changes each
        // pixel through its own tile only, so it fuses into the store. A tile
        // of a light's range takes the light when one of its pixels lies
        // inside the disc and behind the light; its pixels behind the light
        // get colour + alpha 255 (1 - a^2), a = 1 - (-d / r^2).
        if (p.sprites) {
            const int32_t bx = (int32_t)(tx * 4), by = (int32_t)(ty * 4);
            for (uint32_t li = 0; li < p.num_lights; li++) {
                const ResolveLight& L = p.lights[li];
                if (!L.sprite || bx < L.tile_x0 || bx >= L.tile_x1 || by < L.tile_y0 || by >= L.tile_y1) continue;
                const float rx = (float)(int32_t)x + L.half_x, ry = (float)(int32_t)y + L.half_y;
                const float d = __builtin_fmaf(rx, rx, ry * ry) - L.r2;
                const bool front = L.depth > depth;
                if (!(__ballot(inside && d < 0.0f && front) & groupMask)) continue;
                const float a = __builtin_fmaf(d, L.inv_r2, 1.0f);
                const int32_t al = min(max(cvt_rne(__builtin_fmaf(-(a * a), 255.0f, 255.0f)), 0), 255);
                if (front) outc = alpha_blend_u8(outc, L.rgb | ((uint32_t)al << 24));
            }
        }
    }
    if (inside) p.out[p.linear ? (size_t)y * W + x : (size_t)t] = outc;
}

// Coefficients of ColormapTurbo (Shading.cpp:249-260): Google's published
// degree-5 polynomial fit of the Turbo colormap, one row per channel,
// constant term first; double literals rounded to float.
__constant__ float kTurbo[18] = {
    0.13572138, 4.61539260,  -42.66032258, 132.13108234, -152.94239396, 59.28637943,
    0.09140261, 2.19418839,  4.84296658,   -14.18503333, 4.27729857,    2.82956604,
    0.10667330, 12.64194608, -60.58204836, 110.36276771, -89.90310912,  27.34824973,
};

// ResolveDebug's overdraw layers (Shading.cpp:759-766) on a framebuffer
// drawn with the overdraw pipeline: ColormapTurbo of pixels / 30
// (OverdrawPixel) or (pixels + helpers * 0.5) / 30 (OverdrawQuad), packed;
// pixels with depth <= 0 get the background checker. The counter word is
// the clear colour's 16-bit halves plus the counts, saturated. Operation
// order of the reference's optimised build: / 30 as * (1/30), and the
// innermost x * c5 folded into s * (c5 * (1/30)) with s the undivided sum;
// the rest of the Horner chain is FMAs.
template <bool Quad>
__global__ void __launch_bounds__(256) k_resolve_overdraw(const unsigned long long* counts, const uint32_t* depth,
                                                          int width, int height, uint32_t clear_color, uint32_t* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t W = (uint32_t)width, H = (uint32_t)height;
    if (t >= W * H) return;
    const uint32_t tile = t >> 4, q = t & 15u;
    uint32_t tx;
    const uint32_t ty = div_small(tile, W / 4, tx);
    const uint32_t x = tx * 4 + (q & 3u), y = ty * 4 + (q >> 2);
    uint32_t outc;
    if (bitsf(depth[t]) <= 0.0f) {
        outc = ((tx ^ ty) & 1u) ? 0xFFA0A0A0u : 0xFFFFFFFFu;
    } else {
        const unsigned long long c = counts[t];
        const uint32_t hi = (uint32_t)min((unsigned long long)(clear_color >> 16) + (c >> 32), 0xFFFFull);
        const uint32_t lo = (uint32_t)min((unsigned long long)(clear_color & 0xFFFFu) + (c & 0xFFFFFFFFull), 0xFFFFull);
        const float pix = (float)hi;
        const float sum = Quad ? pix + (float)lo * 0.5f : pix;
        const float r = 1.0f / 30.0f;
        const float xs = sum * r;
        float rgb[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float* cf = kTurbo + 6 * k;
            float v = __builtin_fmaf(sum, cf[5] * r, cf[4]);
            v = __builtin_fmaf(xs, v, cf[3]);
            v = __builtin_fmaf(xs, v, cf[2]);
            v = __builtin_fmaf(xs, v, cf[1]);
            rgb[k] = __builtin_fmaf(xs, v, cf[0]);
        }
        outc = pack_rgba8(rgb[0], rgb[1], rgb[2], 1.0f);
    }
    out[(size_t)y * W + x] = outc;
}

}  // namespace

hipError_t launch_resolve_overdraw(const unsigned long long* counts, const uint32_t* depth, int width, int height,
                                   uint32_t clear_color, int quad, uint32_t* out, hipStream_t s, Profiler* prof) {
    const uint32_t n = (uint32_t)width * (uint32_t)height;
    if (quad)
        launch_kernel(prof, "resolve_overdraw_quad", k_resolve_overdraw<true>, dim3((n + 255) / 256), dim3(256), 0, s,
                      counts, depth, width, height, clear_color, out);
    else
        launch_kernel(prof, "resolve_overdraw_pixel", k_resolve_overdraw<false>, dim3((n + 255) / 256), dim3(256), 0,
                      s, counts, depth, width, height, clear_color, out);
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveParams& p, hipStream_t s, Profiler* prof) {
    uint32_t n = (uint32_t)p.width * (uint32_t)p.height;
    static const char* const names[6] = { "resolve_albedo", "resolve_lit", "resolve_normals",
                                          "resolve_metallic_roughness", "resolve_meshlet_id", "resolve_triangle_id" };
    if (p.mode < 0 || p.mode > 5) return hipErrorInvalidValue;
    const dim3 grid((n + 255) / 256), block(256);
    const char* name = names[p.mode];
    switch (p.mode) {
    case 0: launch_kernel(prof, name, k_resolve<0>, grid, block, 0, s, p); break;
    case 1: launch_kernel(prof, name, k_resolve<1>, grid, block, 0, s, p); break;
    case 2: launch_kernel(prof, name, k_resolve<2>, grid, block, 0, s, p); break;
    case 3: launch_kernel(prof, name, k_resolve<3>, grid, block, 0, s, p); break;
    case 4: launch_kernel(prof, name, k_resolve<4>, grid, block, 0, s, p); break;
    default: launch_kernel(prof, name, k_resolve<5>, grid, block, 0, s, p); break;
    }
    return hipGetLastError();
}

}  // namespace glimp
