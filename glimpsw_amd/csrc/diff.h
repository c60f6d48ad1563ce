// Differentiable texture-sampling resolve (diff.hip): forward, backward and
// the workspace of the deterministic segmented reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace glimp {

struct DiffTex {                   // fp32 RGBA mip chain, level m row-major at level_offset[m]
    uint32_t width, height, levels, pad;
    uint64_t level_offset[16];     // in texels
};

// One framebuffer of a differentiable resolve. Pixels of all views are
// numbered consecutively (view v owns [pixel_base, pixel_base + W*H)); a
// view's outputs and gradient inputs are row-major at its pixel base.
struct DiffView {
    const unsigned long long* keys;
    int width, height;
    unsigned long long clear_key;
    const DrawEntry* draws;
    uint32_t num_draws;
    uint32_t pixel_base;
    float object_to_clip[16];
    float scale_x, scale_y, center_x, center_y, kx, ky;
};

struct DiffParams {
    DevMeshlets m;
    const DevMaterial* materials;
    uint32_t num_materials;
    const DiffTex* textures;
    uint32_t num_textures;
    ApproxTables tables;
    uint32_t num_views;
    const DiffView* views;         // device table, or null: `view` holds the only view
    DiffView view;
    uint32_t total_pixels;
    const float4* texels;
    const float2* uv;              // [meshlets * 64]
    const float4* color;           // [meshlets * 64]
    float4* out;                   // forward: [total_pixels]
    const float4* grad_out;        // backward: [total_pixels]
    // backward records, 4 texel and 3 vertex records per pixel slot
    uint32_t* tex_keys;
    float4* tex_vals;
    uint32_t tex_sentinel;         // key of an unused record, above every texel index
    uint32_t* vtx_keys;
    float4* vtx_vals;              // 2 per record: colour gradient, (u, v, 0, 0) gradient
    uint32_t vtx_sentinel;
};

// Device buffers of one backward pass (owned by the context).
struct DiffWorkspace {
    size_t pixels = 0, texels = 0;
    uint32_t* tex_keys = nullptr;
    float4* tex_vals = nullptr;
    uint32_t* vtx_keys = nullptr;
    float4* vtx_vals = nullptr;
    uint32_t* sorted_keys = nullptr;
    uint32_t* sorted_idx = nullptr;
    float4* part_first = nullptr;  // per 256-record tile, up to 2 float4 each
    float4* part_last = nullptr;
    uint32_t* key_last = nullptr;
    uint32_t* spans = nullptr;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
};

hipError_t diff_workspace_reserve(DiffWorkspace& w, size_t pixels);
void diff_workspace_release(DiffWorkspace& w);
hipError_t launch_diff_forward(const DiffParams& p, hipStream_t s, Profiler* prof);
// grad_texels [num_texels], grad_uv [vertices], grad_color [vertices] are
// overwritten (zeros where nothing contributes).
hipError_t launch_diff_backward(DiffParams p, DiffWorkspace& w, uint64_t num_texels, uint32_t num_vertices,
                                float4* grad_texels, float2* grad_uv, float4* grad_color, hipStream_t s,
                                Profiler* prof);

}  // namespace glimp
