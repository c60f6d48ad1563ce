/*
 * CPU oracle for the meshlet raster -> vis-buffer -> resolve path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar restatement of the reference
 * renderer's algorithm (dubiousconst282/GLimpSW, src/SwRast), written from
 * its source text and from probes of how clang lowers its floating-point
 * expressions under the reference's -ffast-math flags. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, as the
 * checker the HIP path is compared against. The product never calls it.
 *
 * Layouts follow the reference:
 *  - meshlets as structure-of-arrays of the Meshlet record (Scene.h:15-30);
 *  - framebuffer layers in 4x4 tiles (Rasterizer.h:10-78);
 *  - textures in TiledY8 layout with 64-texel aligned mip offsets
 *    (Texture.h:482-501, 599-636).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_meshlets {
    uint32_t count;
    const float* bounds;              /* [count][4]  BoundCenter.xyz, BoundRadius */
    const uint8_t* num_vertices;      /* [count] */
    const uint8_t* num_triangles;     /* [count] */
    const uint32_t* material_id;      /* [count], UINT32_MAX = none */
    const uint64_t* tangent_handedness; /* [count] */
    const float* positions;           /* [count][3][64] */
    const uint32_t* texcoords;        /* [count][64] half2 */
    const uint32_t* normal_tangents;  /* [count][64] oct8 x4 */
    const uint8_t* indices;           /* [count][3][128] */
} orc_meshlets;

typedef struct orc_texture {
    uint32_t width, height, mip_levels, num_layers;
    uint32_t row_shift, layer_stride;
    float scale_u, scale_v, scale_lerp_u, scale_lerp_v;
    int32_t mask_u, mask_v, mask_lerp_u, mask_lerp_v;
    uint32_t mip_offsets[16];
    const uint32_t* data;             /* layer_stride * num_layers texels (+ 64 slack) */
} orc_texture;

typedef struct orc_material {
    int32_t texture;                  /* index into textures, -1 = none */
    uint8_t double_sided;
    uint8_t alpha_cutoff;
    uint8_t pad[2];
} orc_material;

typedef struct orc_shading_ctx {
    const orc_meshlets* meshlets;
    const orc_material* materials;
    const orc_texture* textures;
    uint32_t meshlet_offset;
    const uint16_t* cull_bitmap;      /* 1 bit per meshlet, 16 per word; NULL = all visible */
    float object_to_clip[16];         /* column-major, m[c*4+r] */
    float object_to_world[9];         /* mat3 column-major (the deferred pipeline's normals) */
} orc_shading_ctx;

typedef struct orc_framebuffer {
    uint32_t width, height, layer_stride, num_layers;
    uint32_t* data;                   /* layer_stride * num_layers */
} orc_framebuffer;

typedef struct orc_raster_opts {
    uint8_t enable_clipping;
    uint8_t enable_guardband;
    uint8_t pipeline;                 /* ORC_PIPELINE_*: the fragment shader of the dispatch table */
    uint8_t pad;
    uint32_t* coverage_count;         /* optional [height][width]: covered samples per pixel, before depth test */
} orc_raster_opts;

/* Dispatch tables of ShadingContext (Shading.cpp:650-656). */
#define ORC_PIPELINE_VISBUFFER 0u     /* FS_EncodeSurfaceId<false/true> */
#define ORC_PIPELINE_DEFERRED 1u      /* FS_EncodeGBuffer (three layers) */
#define ORC_PIPELINE_OVERDRAW 2u      /* FS_Overdraw */

typedef struct orc_counters {
    uint64_t triangles_processed;
    uint64_t triangles_rasterized;
    uint64_t triangles_clipped;
    uint64_t fragments_covered;       /* covered samples before the depth test */
    uint64_t fragments_passed;        /* samples that passed the depth test */
} orc_counters;

typedef struct orc_light {
    uint32_t type;                    /* 0 directional, 1 point, 2 spot */
    float position[3];
    float direction[3];
    float color[3];
    float intensity;
    float inv_radius_sq, spot_scale, spot_offset;
} orc_light;

typedef struct orc_resolve_params {
    float inv_screen_proj[16];        /* GetInverseScreenProjMatrix(WorldToClipMat, size) (Camera.h:140-146) */
    float object_to_clip[16];
    float object_to_world[9];         /* mat3 column-major */
    float view_pos[3];
    float exposure;
    const orc_light* lights;
    uint32_t num_lights;
    uint32_t light_sprites;           /* draw point/spot lights over the image (Shading.cpp:689-727) */
    float world_to_clip[16];          /* ShadingContext::WorldToClipMat (light sprites) */
    const orc_texture* skybox;        /* ShadingContext::SkyboxTex: R11G11B10f octahedral map, or NULL */
} orc_resolve_params;

/* Tables reproducing VRCP14PS / VRSQRT14PS; must be loaded before any call. */
void orc_load_tables(const uint16_t* tables /* 131072 entries */);
float orc_rcp14(float x);
float orc_rsqrt14(float x);
float orc_div(float a, float b);      /* vector a/b under -ffast-math */
int32_t orc_cvt_rne(float x);         /* cvtps2dq */
/* op 0: rcp14(a), 1: rsqrt14(a), 2: fast-math a / b, over n elements. */
void orc_approx_n(int32_t op, const float* a, const float* b, float* out, uint64_t n);

void orc_fb_clear(orc_framebuffer* fb, uint32_t color, float depth);
void orc_get_pixels(const orc_framebuffer* fb, uint32_t layer, uint32_t* dest, uint32_t stride);
void orc_draw_meshlets(orc_framebuffer* fb, uint32_t count, const orc_shading_ctx* ctx,
                       const orc_raster_opts* opts, orc_counters* counters);
/* The same draw restricted to the framebuffer rows [row0, row1) (multiples
 * of 4): every meshlet is shaded, set up and clipped, and only fragments in
 * those rows are drawn and counted. Calls on disjoint row ranges may run on
 * separate threads and together give the whole frame bit for bit. */
void orc_draw_meshlets_rows(orc_framebuffer* fb, uint32_t count, const orc_shading_ctx* ctx,
                            const orc_raster_opts* opts, orc_counters* counters, uint32_t row0, uint32_t row1);

/* Frustum cull of bound spheres against 5 normalised planes (xyz,w each). */
/* Occlusion inputs of CullMeshlets (Shading.cpp:775-869): the previous
 * frame's depth map (levels 0 .. mip_levels - 2 of orc_hiz_build, row-major,
 * level m of ((W-1) >> (m+1)) + 1 columns), objectToPrevViewMat
 * (column-major), objectToWorldScale, projFactors (znear, f/aspect, -f) and
 * the frame size. */
typedef struct orc_occlusion {
    const float* depth_map;
    uint32_t mip_levels;
    float object_to_prev_view[16];
    float object_to_world_scale;
    float proj[3];
    float frame_size[2];
} orc_occlusion;
uint32_t orc_hiz_levels(uint32_t width, uint32_t height);
void orc_hiz_build(const orc_framebuffer* fb, float* out);
uint32_t orc_cull_meshlets_occlusion(uint16_t* bitmap, const orc_meshlets* m, uint32_t first, uint32_t count,
                                     const float planes[20], const orc_occlusion* occ);
void orc_hiz_sphere16(const float* sphere, const orc_occlusion* o, int32_t* mask, float* bb, int32_t* mip,
                      int32_t* rect, float* depth);
uint32_t orc_cull_meshlets(uint16_t* bitmap, const orc_meshlets* m, uint32_t first, uint32_t count,
                           const float planes[20]);

/* ResolveDebug(BaseColor): layer 0 becomes RGBA8 albedo, sky = checker bg. */
void orc_resolve_albedo(orc_framebuffer* fb, const orc_shading_ctx* ctx);
/* ResolveDebug (Shading.cpp:734-773): layer ORC_DEBUG_*; object_to_world
 * (mat3) for the Normals and MetallicRoughness layers. */
#define ORC_DEBUG_BASE_COLOR 0u
#define ORC_DEBUG_NORMALS 1u
#define ORC_DEBUG_METALLIC_ROUGHNESS 2u
#define ORC_DEBUG_MESHLET_ID 3u
#define ORC_DEBUG_TRIANGLE_ID 4u
#define ORC_DEBUG_OVERDRAW_PIXEL 5u   /* framebuffer drawn with ORC_PIPELINE_OVERDRAW */
#define ORC_DEBUG_OVERDRAW_QUAD 6u
void orc_resolve_debug(orc_framebuffer* fb, const orc_shading_ctx* ctx, uint32_t layer, const float* object_to_world);
/* The skybox lookup of Resolve (Shading.cpp:675-678) for 16 lanes:
 * texel coordinates of SampleLevel from view directions dir [3][16]
 * (ScaleLerpU/V slu, slv), and the R11G11B10f bilinear filter of four
 * texels d [4][16] (00, 10, 01, 11) at fixed-point coordinates ixf, iyf
 * with the x-inbound mask inx: out = rgb [3][16]. */
void orc_sky_coords16(const float* dir, float slu, float slv, int32_t* ix, int32_t* iy);
void orc_env_bilinear16(const uint32_t* d, const int32_t* ixf, const int32_t* iyf, const int32_t* inx, float* out);
/* FS_EncodeGBuffer's layer-2 word for 16 lanes: n = remapped world normal
 * [3][16], nmr = layer-1 texels. */
void orc_gbuffer_pack16(const float* n, const uint32_t* nmr, uint32_t* out);
/* The overdraw layers' colour for 16 counter words: out = rgb [3][16]. */
void orc_overdraw_colour16(const uint32_t* data, int32_t quad, float* out);
/* Resolve(): lit, tonemapped RGBA8 into layer 0. */
void orc_resolve_lit(orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_resolve_params* p);
/* The lit resolve of rows [row0, row1) only (multiples of 4), light sprites
 * included; calls on disjoint row ranges may run on separate threads. */
void orc_resolve_lit_rows(orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_resolve_params* p, uint32_t row0,
                          uint32_t row1);
/* The lit colour of one 4x4 tile at given surfaces (Shading.cpp:602-645,
 * 221-226, 680): albedo and layer-1 (normal/metallic/roughness) texels,
 * world normals [3][16] and world positions [3][16] of the 16 lanes, lanes
 * with active[l] != 0 lit, from p's lights, view position and exposure; out =
 * Tonemap_Unreal(colour * exposure) packed to RGBA8 (inactive lanes 0). */
void orc_eval_lighting16(const uint32_t* albedo, const uint32_t* nmr, const float* normal, const float* wpos,
                         const uint8_t* active, const orc_resolve_params* p, uint32_t* out);
/* RGBA8u::UnpackSrgb (Texture.h:37-54) of n texels: out [n][4]. */
void orc_unpack_srgb_n(const uint32_t* packed, float* out, uint64_t n);

/* Sampler entry for tests: 16 lanes = one 4x4 tile (Texture.h:412-459). */
void orc_sample_level16(const orc_texture* tex, const float* u, const float* v, int32_t layer,
                        const int32_t* mip, int32_t wrap, int32_t mag_linear, int32_t min_linear,
                        uint32_t* out);
void orc_intersect_triangle16(const float* clip /*[3][4][16]*/, const float* suv /*[2][16]*/,
                              float sw, float sh, float* bary, float* ddx, float* ddy);
void orc_edge16(const int32_t* det, const float* Z /*[3][16]*/, float* Z10, float* Z20);
void orc_uvgrad16(const float* t /*[3][2][16]*/, const float* bary, const float* ddx, const float* ddy,
                  float su, float sv, float* uv /*[2][16]*/, int32_t* mip);
int32_t orc_clip_triangle(const float* v /*[3][4]*/, uint32_t outcodes, float gbx, float gby,
                          float* out /*[128][3][4]*/);
void orc_setup16(const float* clip /*[3][4][16]*/, int32_t w, int32_t h, int32_t cull_mode,
                 int32_t* X, int32_t* Y, float* Z, float* W, float* det, uint16_t* mask);

/* Per-fragment depth and perspective-correct barycentrics of 16 fragments
 * (Rasterizer.h:293-319) from the edge values E1, E2 and
 * coef = {Z0, Z10, Z20, W0, W0S, W1S, W2S}. */
void orc_fragment16(const int32_t* e1, const int32_t* e2, const float* coef, float* depth, float* bu, float* bv);
/* Normal mapping of the lit resolve (Shading.cpp:556-572) for 16 lanes:
 * n, t = world-space normal and tangent [3][16], hand = 0 or 1<<31,
 * nmr = layer-1 texels; out [3][16]. */
void orc_normal_map16(const float* n, const float* t, const uint32_t* hand, const uint32_t* nmr, float* out);
/* Light sprite pieces (Shading.cpp:692-724): placement (out = depth, centre
 * x, y, radius; tiles = x0, y0, x1, y1; returns 0 when culled), one tile's
 * distances and colours, and AlphaBlendU8 (Shading.cpp:238-246). */
int32_t orc_sprite_place(const float* m, const float* pos, uint32_t W, uint32_t H, float* out, int32_t* tiles);
void orc_sprite_tile(int32_t x, int32_t y, float cx, float cy, float radius, const float* color, float* dist,
                     uint32_t* fg);
uint32_t orc_alpha_blend_u8(uint32_t bg, uint32_t fg);
/* The alpha test's UVs and implicit mip levels for one 4x4 tile
 * (FS_EncodeSurfaceId<true>, Shading.cpp:316-327; SampleImplicitLod,
 * Texture.h:403-410): remap = {ClippedU[3], ClippedV[3]} or NULL,
 * uv = vertex UVs [3][2], slu/slv = ScaleLerpU/V. */
void orc_alpha_uv16(const int32_t* e1, const int32_t* e2, const float* coef, const float* remap, const float* uv,
                    float slu, float slv, float* U, float* V, int32_t* mip);

/* ------------------------------------------------------------ differentiable resolve
 *
 * Not a reference operation: the texture-sampling resolve made
 * differentiable. For a covered pixel of surface (meshlet, triangle), with
 * barycentrics b and vertex parameters uv_k (fp32 [meshlets][64][2]) and
 * color_k (fp32 [meshlets][64][4]):
 *   U,V   = uv_0 + b1 (uv_1 - uv_0) + b2 (uv_2 - uv_0)
 *   col   = color_0 + b1 (color_1 - color_0) + b2 (color_2 - color_0)
 *   out   = Sample(texture of the material, U, V) * col      (RGBA)
 * Sample follows SampleLevel's addressing (Texture.h:412-459) with float
 * weights: Repeat wrap, level and filter chosen as the resolve does (mip from
 * the UV gradients, magnification anywhere in the 4x4 tile -> bilinear at
 * level 0 with texel centres at +0.5 and clamped edges, else nearest texel
 * of level clamp(mip)). Texels are fp32 RGBA, level m of a texture stored
 * row-major at level_offset[m]. Uncovered pixels give 0; untextured
 * materials sample 1.
 *
 * Barycentrics come from IntersectTriangle's formula with correctly rounded
 * divisions, in fp32 as the GPU evaluates them, and are constants of the
 * function, like coverage (orc_bary_f64 checks that formula against the
 * textbook one in double).
 *
 * The oracle takes every discrete choice (surface, level, filter, wrap
 * period, texel cell, edge clamps) from the fp32 evaluation the GPU performs
 * and evaluates everything continuous in double, so its gradient is the FP64
 * gradient of the same piecewise function. */
typedef struct orc_diff_texture {
    uint32_t width, height, mip_levels, pad;
    uint64_t level_offset[16];        /* in texels (4 floats each) */
} orc_diff_texture;

typedef struct orc_diff_params {
    float object_to_clip[16];
    const orc_diff_texture* textures; /* indexed by material texture */
    uint32_t num_textures;
    const float* texels;              /* fp32 params: decisions are taken from these */
    const float* uv;
    const float* color;
    const double* texels_d;           /* optional double values of the same params */
    const double* uv_d;               /* (finite differences); NULL = use the fp32 ones */
    const double* color_d;
} orc_diff_params;

/* out: [height][width][4] double, row-major. */
void orc_diff_forward(const orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_diff_params* p, double* out);
/* Gradients of sum(grad_out * out) w.r.t. texels, uv and color (dense,
 * accumulated into the given zeroed arrays). */
void orc_diff_backward(const orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_diff_params* p,
                       const double* grad_out, double* grad_texels, double* grad_uv, double* grad_color);
/* Perspective-correct barycentrics of one pixel centre in double from
 * textbook formulas (for checking the fp32 IntersectTriangle). */
void orc_bary_f64(const float* clip /*[3][4]*/, double su, double sv, double* bary);

#ifdef __cplusplus
}
#endif
