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
};

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
    ResolveLight lights[kMaxResolveLights];
    int mode;                      // 0 albedo (ResolveDebug BaseColor), 1 lit (Resolve)
    uint32_t* out;                 // RGBA8 per pixel
    int linear;
};

hipError_t launch_resolve(const ResolveParams& p, hipStream_t s, Profiler* prof);

}  // namespace glimp
