// Resolve pass parameters (resolve.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace glimp {

constexpr int kMaxResolveLights = 8;

struct ResolveLight {              // Light (Scene.h:52-75), precomputed terms included
    uint32_t type;                 // 0 directional, 1 point, 2 spot
    float position[3];
    float direction[3];
    float color[3];
    float intensity;
    float inv_radius_sq, spot_scale, spot_offset;
    // its sprite (Shading.cpp:692-724), placed on the host; sprite = 0: none
    uint32_t sprite;
    int32_t tile_x0, tile_y0, tile_x1, tile_y1;   // tile bases x0 <= x < x1, y0 <= y < y1
    float depth;                   // the light's depth
    float half_x, half_y;          // 0.5 - centre
    float r2, inv_r2;              // radius^2 and 1 / radius^2
    uint32_t rgb;                  // RGBA8u::Pack of the colour, alpha byte 0
};
constexpr size_t kLightAbiBytes = 56;   // the glimp_light prefix

struct ResolveParams {
    DevMeshlets m;
    const DevMaterial* materials;
    uint32_t num_materials;
    const DevTexture* textures;
    ApproxTables tables;
    const unsigned long long* keys;
    int width, height;
    unsigned long long clear_key;
    const DrawEntry* draws;
    uint32_t num_draws;
    float object_to_clip[16];
    float object_to_world[9];
    float inv_screen_proj[16];
    float view_pos[3];
    float exposure, light_exposure;
    float scale_x, scale_y, center_x, center_y, kx, ky;  // screen UV and derivative steps
    uint32_t num_lights;
    uint32_t sprites;              // a light has a sprite
    ResolveLight lights[kMaxResolveLights];
    DevTexture sky;                // ShadingContext::SkyboxTex (R11G11B10f), when has_sky
    uint32_t has_sky;
    int mode;                      // GLIMP_RESOLVE_*: 0 BaseColor, 1 lit (Resolve), 2-5 other debug layers
    uint32_t* out;                 // RGBA8 per pixel
    int linear;
};

hipError_t launch_resolve(const ResolveParams& p, hipStream_t s, Profiler* prof);
// ResolveDebug(OverdrawPixel / OverdrawQuad) of an overdraw framebuffer
// (counts and depth as launch_draw_overdraw leaves them) into row-major RGBA8.
hipError_t launch_resolve_overdraw(const unsigned long long* counts, const uint32_t* depth, int width, int height,
                                   uint32_t clear_color, int quad, uint32_t* out, hipStream_t s, Profiler* prof);

}  // namespace glimp
