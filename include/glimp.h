/*
 * glimp: C ABI of the MI355X meshlet raster -> vis-buffer -> resolve path.
 *
 * This is the boundary a host renderer binds to. Each entry point replaces
 * one operation of the reference renderer (dubiousconst282/GLimpSW,
 * src/SwRast); the reference location is cited on each declaration.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative GLIMP_E_* code on
 *    failure; glimp_last_error() then describes the failure (thread local).
 *  - No exceptions cross this boundary. Plain pointers and sizes only.
 *  - Pointers documented as "host" are read or written synchronously before
 *    the call returns. Pointers documented as "device" must be device
 *    memory of the context's GPU.
 *  - `stream` is a hipStream_t (NULL = the default stream). Work is
 *    enqueued asynchronously unless a function says otherwise.
 *  - A context is bound to one GPU; use one context per host thread.
 */
#ifndef GLIMP_H
#define GLIMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLIMP_OK 0
#define GLIMP_E_INVALID (-1)      /* bad argument */
#define GLIMP_E_HIP (-2)          /* HIP runtime error */
#define GLIMP_E_NOMEM (-3)        /* allocation failed */
#define GLIMP_E_LIMIT (-4)        /* a size limit was exceeded */
#define GLIMP_E_UNSUPPORTED (-5)  /* feature not built */

/* Largest framebuffer side: the 32-bit edge-equation bound
 * (Rasterizer.h:203, MaxRenderSize). */
#define GLIMP_MAX_RENDER_SIZE 2896u
/* Meshlet limits (Rasterizer.h:83, ShadedMeshlet::MaxVertices/MaxPrims). */
#define GLIMP_MAX_VERTICES 64u
#define GLIMP_MAX_PRIMS 128u
/* Meshlets that may be drawn into one framebuffer between two clears:
 * the depth test key carries a 21-bit draw sequence number. */
#define GLIMP_MAX_DRAWN_MESHLETS (1u << 21)

typedef struct glimp_context glimp_context;
typedef struct glimp_scene glimp_scene;
typedef struct glimp_texture glimp_texture;
typedef struct glimp_framebuffer glimp_framebuffer;

/* ------------------------------------------------------------ context */

/* Creates a context on HIP device `device`. No reference counterpart: the
 * reference's Rasterizer(numThreads) (Rasterizer.h:210) owns a CPU thread
 * pool; here the context owns device scratch and the approximation tables. */
int glimp_create(glimp_context** out, int device);
void glimp_destroy(glimp_context* ctx);
/* Message for the last failure on this thread ("" if none). */
const char* glimp_last_error(void);

/* ------------------------------------------------------------ resources */

/* Meshlets as structure-of-arrays of the reference Meshlet record
 * (Scene.h:15-30); all host pointers, `count` entries each. */
typedef struct glimp_meshlets_desc {
    uint32_t count;
    const float* bounds;               /* [count][4] BoundCenter.xyz, BoundRadius */
    const uint8_t* num_vertices;       /* [count] (<= 64) */
    const uint8_t* num_triangles;      /* [count] (<= 128) */
    const uint32_t* material_id;       /* [count], 0xFFFFFFFF = none */
    const uint64_t* tangent_handedness;/* [count] */
    const float* positions;            /* [count][3][64] */
    const uint32_t* texcoords;         /* [count][64] half2 */
    const uint32_t* normal_tangents;   /* [count][64] octahedral n.xy t.xy, unorm8 */
    const uint8_t* indices;            /* [count][3][128] */
} glimp_meshlets_desc;

/* Uploads meshlets (the reference keeps them in Scene::Meshlets,
 * Scene.h:117-128). */
int glimp_scene_create(glimp_context* ctx, const glimp_meshlets_desc* desc, glimp_scene** out);
void glimp_scene_destroy(glimp_scene* scene);

/* RGBA8 texture storage exactly as CreateTexture2D lays it out
 * (Texture.h:599-636): TiledY8 texels, mip offsets in texels, layers
 * `layer_stride` texels apart. `texels` (host) holds
 * layer_stride * num_layers entries. */
#define GLIMP_TEXEL_RGBA8 0u          /* pixfmt::RGBA8u (Texture.h:24-68): material textures */
#define GLIMP_TEXEL_R11G11B10F 1u     /* pixfmt::R11G11B10f (Texture.h:135-185): HdrTexture2D, the skybox */
typedef struct glimp_texture_desc {
    uint32_t width, height, mip_levels, num_layers;
    uint32_t row_shift, layer_stride;
    uint32_t mip_offsets[16];
    const uint32_t* texels;
    uint32_t format;                   /* GLIMP_TEXEL_* */
} glimp_texture_desc;
int glimp_texture_create(glimp_context* ctx, const glimp_texture_desc* desc, glimp_texture** out);
void glimp_texture_destroy(glimp_texture* tex);

/* Material table (Scene.h:7-14). `texture` indexes the `textures` array
 * given with it (-1 = none). */
typedef struct glimp_material {
    int32_t texture;
    uint8_t double_sided;
    uint8_t alpha_cutoff;
    uint8_t pad[2];
} glimp_material;
int glimp_scene_set_materials(glimp_scene* scene, const glimp_material* materials, uint32_t count,
                              glimp_texture* const* textures, uint32_t num_textures);

/* Vis-buffer: per pixel one 64-bit depth-test key in 4x4 tiles, the device
 * form of the reference's two-layer Framebuffer (Rasterizer.h:10-78).
 * Width and height must be multiples of 4 and <= GLIMP_MAX_RENDER_SIZE. */
int glimp_framebuffer_create(glimp_context* ctx, uint32_t width, uint32_t height, glimp_framebuffer** out);
void glimp_framebuffer_destroy(glimp_framebuffer* fb);

/* Framebuffer::Clear(color, depth) (Rasterizer.h:35-38). `depth` must be
 * >= 0 (reverse-Z; fragments pass the test only when strictly nearer). */
int glimp_framebuffer_clear(glimp_framebuffer* fb, uint32_t color, float depth, void* stream);

/* Framebuffer::GetPixels for the id layer and the depth layer
 * (ImageHelpers.cpp:109-147). Writes row-major (linear != 0) or 4x4-tiled
 * (linear == 0) host arrays of width*height entries; either may be NULL.
 * Synchronous. */
int glimp_framebuffer_read(glimp_framebuffer* fb, uint32_t* ids, float* depth, int linear, void* stream);

/* Framebuffer::GetPixels of one layer (ImageHelpers.cpp:109-147) as 32-bit
 * words: layer 0 = surface ids (vis-buffer), counters (overdraw) or base
 * colour (deferred); layer 1 = depth bits; layer 2 = the deferred
 * pipeline's packed normal / metallic / roughness. Row-major (linear != 0)
 * or 4x4-tiled host array of width*height words. Synchronous. */
int glimp_framebuffer_read_layer(glimp_framebuffer* fb, uint32_t layer, uint32_t* out, int linear, void* stream);

/* Device pointer to the raw 64-bit key buffer (4x4 tiles) and its size. */
int glimp_framebuffer_keys(glimp_framebuffer* fb, void** keys, size_t* bytes);

/* ------------------------------------------------------------ drawing */

#define GLIMP_DRAW_NO_CLIPPING 0x1u   /* Rasterizer::EnableClipping = false (Rasterizer.h:207) */
#define GLIMP_DRAW_NO_GUARDBAND 0x2u  /* Rasterizer::EnableGuardband = false (Rasterizer.h:208) */
#define GLIMP_DRAW_DIRECT 0x4u        /* per-fragment global atomics instead of screen bins */
#define GLIMP_DRAW_COUNTERS 0x8u       /* maintain the perf counters (adds atomics) */

/* Pre-instantiated shader pipelines: the device counterpart of a dispatch
 * table built by GetDispatchTable<MeshFn, FragFn...> (Rasterizer.h:342-364). */
#define GLIMP_PIPELINE_VISBUFFER 0u   /* ShadingContext::VisBufferShader (Shading.cpp:650-654) */
#define GLIMP_PIPELINE_DEFERRED 1u    /* ShadingContext::DeferredShader: FS_EncodeGBuffer (Shading.cpp:344-407, 655).
                                       * Layer 0 holds the base colour, layer 1 the depth, layer 2 the
                                       * octahedral normal (2 x 10 bits) and the metallic / roughness bytes
                                       * cut to 6 bits; a clear leaves layer 2 alone. Every drawn meshlet
                                       * needs a material with a texture (GLIMP_E_UNSUPPORTED otherwise).
                                       * Always rasterized with per-fragment atomics (GLIMP_DRAW_DIRECT). */
#define GLIMP_PIPELINE_OVERDRAW 2u    /* ShadingContext::OverdrawShader: FS_Overdraw (Shading.cpp:333-342, 656).
                                       * Layer 0 holds per pixel the covered count (high 16 bits) and the
                                       * quad-helper count (low 16 bits) added to the clear colour
                                       * (saturating), layer 1 the maximum plane depth of the tiles drawn.
                                       * Always rasterized with per-fragment atomics (GLIMP_DRAW_DIRECT). */

typedef struct glimp_draw_desc {
    uint32_t meshlet_offset;           /* ShadingContext::MeshletOffset */
    uint32_t meshlet_count;
    float object_to_clip[16];          /* ShadingContext::ObjectToClipMat, column-major */
    const float* frustum_planes;       /* host [5][4] normalised planes, or NULL: no culling */
    uint32_t flags;                    /* GLIMP_DRAW_* */
    uint32_t pipeline;                 /* GLIMP_PIPELINE_* */
    uint32_t bin_size;                 /* screen bin side in pixels: 8, 16, 32 or 64; 0 = default */
    uint32_t bin_capacity;             /* records per bin before spilling, 0 = default */
    const struct glimp_occlusion* occlusion; /* with frustum_planes: HiZ test of the survivors, or NULL */
    float object_to_world[9];          /* ShadingContext::ObjectToWorldMat, mat3 column-major (deferred) */
} glimp_draw_desc;

/* Rasterizer::DrawMeshlets(fb, count, {VisBufferShader, &ctx})
 * (Rasterizer.h:213), preceded by ShadingContext::CullMeshlets' frustum
 * test (Shading.cpp:775-869) when planes are given. Depth-test ties resolve
 * in the reference's single-thread draw order. A framebuffer holds the
 * output of one pipeline between clears: drawing with another pipeline
 * before clearing fails with GLIMP_E_INVALID. Asynchronous. */
int glimp_draw_meshlets(glimp_context* ctx, glimp_framebuffer* fb, glimp_scene* scene,
                        const glimp_draw_desc* desc, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same draw as glimp_draw_meshlets with the
 * vis-buffer pipeline on the direct route (GLIMP_DRAW_DIRECT: bin_size and
 * bin_capacity are ignored), with the same results, but every scratch buffer
 * comes from a static device array of about 16 MB that each GPU holds for the
 * life of the process: the call makes no allocation. Limits: at most 1024
 * meshlets per call (GLIMP_E_LIMIT beyond) and 256 draws per framebuffer
 * between clears. All calls on one GPU share that GPU's array, so calls that
 * overlap on one GPU (other host threads, or other streams without ordering
 * between them) race; order them on one stream. Asynchronous. */
int glimp_draw_meshlets_static(glimp_context* ctx, glimp_framebuffer* fb, glimp_scene* scene,
                               const glimp_draw_desc* desc, void* stream);

/* ShadingContext::CullMeshlets frustum branch (Shading.cpp:775-869): writes
 * one bit per meshlet, 16 per word (host `bitmap`), returns the visible
 * count in *visible. Synchronous. */
int glimp_cull_meshlets(glimp_context* ctx, glimp_scene* scene, uint32_t first, uint32_t count,
                        const float planes[20], uint16_t* bitmap, uint32_t* visible, void* stream);

/* The previous frame's depth map for occlusion culling: a min-depth pyramid
 * of a width x height framebuffer with the level count of
 * CreateTexture2D<R32f>(halfW, halfH, 16) (Main.cpp:54-56, Texture.h:599-636);
 * level m texel covers 2^(m+1) x 2^(m+1) pixels. */
typedef struct glimp_hiz glimp_hiz;
int glimp_hiz_create(glimp_context* ctx, uint32_t width, uint32_t height, glimp_hiz** out);
void glimp_hiz_destroy(glimp_hiz* hiz);
/* texutil::DownsampleDepth (ImageHelpers.cpp:241-245): rebuilds the pyramid
 * from fb's depth (same size as the map). Asynchronous. */
int glimp_hiz_build(glimp_hiz* hiz, glimp_framebuffer* fb, void* stream);
/* Copies the sampled levels (0 .. levels - 2, row-major, level m with
 * ((width-1) >> (m+1)) + 1 columns, concatenated) to host `out` (NULL: only
 * the sizes); *levels = the map's level count, *floats = the total.
 * Synchronous. */
int glimp_hiz_read(glimp_hiz* hiz, float* out, uint32_t* levels, uint64_t* floats, void* stream);

/* Occlusion inputs of CullMeshlets (Shading.cpp:775-869): the depth map of
 * the previous frame and the terms the reference derives from its matrices:
 * objectToPrevViewMat = prevViewMat * modelMat, objectToWorldScale =
 * length(modelMat[0]), projFactors = (projMat[3][2], projMat[0][0],
 * projMat[1][1]). The frame size is the depth map's. */
typedef struct glimp_occlusion {
    const glimp_hiz* depth_map;
    float object_to_prev_view[16];     /* column-major */
    float object_to_world_scale;
    float proj_factors[3];             /* znear, f/aspect, -f */
} glimp_occlusion;

/* CullMeshlets with the occlusion branch (Shading.cpp:811-840): frustum test,
 * then the HiZ test of the survivors. As glimp_cull_meshlets. */
int glimp_cull_meshlets_occlusion(glimp_context* ctx, glimp_scene* scene, uint32_t first, uint32_t count,
                                  const float planes[20], const glimp_occlusion* occlusion, uint16_t* bitmap,
                                  uint32_t* visible, void* stream);

/* ------------------------------------------------------------ import
 *
 * Host-side helpers of Scene::ImportGltf (Scene.cpp:199-279); no device
 * work. The reference uses meshoptimizer v1.1 there, which this library
 * does not contain: glimp_build_meshlets is its own greedy builder with the
 * same limits and output layout (meshopt_buildMeshlets, Scene.cpp:214-217),
 * so meshlet partitions differ from the reference's. */
typedef struct glimp_meshlet_range {
    uint32_t vertex_offset;            /* into meshlet_vertices */
    uint32_t triangle_offset;          /* into meshlet_triangles (3 u8 corners per triangle) */
    uint32_t vertex_count, triangle_count;
} glimp_meshlet_range;
/* meshopt_buildMeshletsBound (Scene.cpp:197): the most meshlets
 * glimp_build_meshlets can write for index_count indices. */
size_t glimp_build_meshlets_bound(size_t index_count, size_t max_vertices, size_t max_triangles);
/* meshopt_buildMeshlets (Scene.cpp:214-217): partitions a triangle list into
 * meshlets of at most max_vertices (<= 256) vertices and max_triangles
 * (<= 512) triangles; `meshlets` holds glimp_build_meshlets_bound entries,
 * meshlet_vertices and meshlet_triangles index_count entries each. Every
 * input triangle lands in exactly one meshlet. */
int glimp_build_meshlets(const uint32_t* indices, size_t index_count, const float* positions, size_t vertex_count,
                         size_t vertex_stride, uint32_t max_vertices, uint32_t max_triangles,
                         glimp_meshlet_range* meshlets, uint32_t* meshlet_vertices, uint8_t* meshlet_triangles,
                         size_t* meshlet_count);
/* meshopt_computeMeshletBounds' sphere (Scene.cpp:226-232): a bounding
 * sphere (centre, radius) of the meshlet's vertices that contains each of
 * them in float arithmetic. */
int glimp_meshlet_bounds(const uint32_t* meshlet_vertices, uint32_t vertex_count, const float* positions,
                         size_t vertex_stride, float sphere[4]);
/* Meshlet::NormalTangents of ImportGltf (Scene.cpp:266-274):
 * packUnorm4x8(MapOctahedron(normal), MapOctahedron(tangent.xyz)) per vertex,
 * the octahedral map evaluated as the reference's optimised build does;
 * tangents ([count][4]) may be NULL (zero bytes). */
int glimp_pack_normal_tangents(const float* normals, const float* tangents, size_t count, uint32_t* packed);

/* ------------------------------------------------------------ batches */

/* Many views (scene, meshlet range, matrix, framebuffer) rendered together:
 * every stage of the pipeline is one launch over all views, and the whole
 * batch can be replayed as a single hipGraph launch. No reference
 * counterpart: the reference renders one view per DrawMeshlets call
 * (Main.cpp:216-245); a batch gives each view exactly the vis-buffer that a
 * clear plus one DrawMeshlets would. */
typedef struct glimp_batch glimp_batch;

typedef struct glimp_view_desc {
    glimp_scene* scene;
    glimp_framebuffer* framebuffer;    /* one framebuffer per view, not shared */
    uint32_t meshlet_offset, meshlet_count;
    float object_to_clip[16];
    const float* frustum_planes;       /* host [5][4] or NULL */
    uint32_t flags;                    /* GLIMP_DRAW_NO_CLIPPING / GLIMP_DRAW_NO_GUARDBAND */
    uint32_t clear_color;
    float clear_depth;                 /* >= 0 */
    const struct glimp_occlusion* occlusion; /* with frustum_planes: HiZ test of the survivors, or NULL;
                                        * copied at creation, its depth map read at every launch */
} glimp_view_desc;

/* Launch shape: screen bin side, records per bin, bin-raster workgroup size. */
typedef struct glimp_batch_shape {
    uint32_t bin_size;                 /* 8, 16, 32 or 64 */
    uint32_t bin_capacity;             /* records per bin before spilling */
    uint32_t raster_threads;           /* 128 or 256; also 512 or 1024 for 32- and 64-pixel bins */
} glimp_batch_shape;

int glimp_batch_create(glimp_context* ctx, const glimp_view_desc* views, uint32_t count,
                       const glimp_batch_shape* shape /* NULL = default */, glimp_batch** out);
void glimp_batch_destroy(glimp_batch* batch);
int glimp_batch_set_shape(glimp_batch* batch, const glimp_batch_shape* shape);
/* Clears each view's framebuffer and draws its meshlets, one launch per
 * stage for all views, then the resolves if set; use_graph != 0 replays the
 * batch as one hipGraph on `stream` (captured on first use, re-captured
 * after a shape or resolve change). Asynchronous. The framebuffers and
 * scenes must outlive the batch. */
int glimp_batch_launch(glimp_batch* batch, int use_graph, void* stream);
/* The same work as glimp_framebuffer_clear + glimp_draw_meshlets (and
 * glimp_resolve when set) per view, view after view (for comparison with
 * the batched launch). */
int glimp_batch_launch_per_view(glimp_batch* batch, void* stream);
/* Autotuner: times the candidate shapes (bin 8-64, capacity 256-4096,
 * 128/256-thread bin raster, 512/1024 for 32/64-pixel bins) on this batch,
 * keeps the fastest and returns it with its time per launch. Bin sizes go
 * 16, 32, 8, 64 and capacities from the largest down; the first launch of a
 * bin size gives its largest bin's record count, and smaller capacities
 * (which would overflow) are not launched; when even the largest overflows,
 * only it is tried (overflowing records take the overflow raster, whose cost
 * the timing includes). Each candidate is probed with one launch first:
 * shapes that drop anything are rejected, one slower than twice the best so
 * far is not timed further, and a bin size whose first launch takes over
 * four times the best is given up. If every shape is rejected the batch's
 * bin size is kept with the largest capacity. Synchronous. */
int glimp_batch_autotune(glimp_batch* batch, void* stream, glimp_batch_shape* chosen, float* best_ms);

/* ------------------------------------------------------------ differentiable resolve */

/* The texture-sampling resolve as a differentiable function of fp32 texel
 * mip chains and per-vertex UVs and colours, with coverage and barycentrics
 * held fixed by the vis-buffer. No reference counterpart: it extends
 * ResolveDebug(BaseColor) (Shading.cpp:734-773) and SampleLevel
 * (Texture.h:412-459). For a covered pixel:
 *   U,V = uv0 + b1 (uv1 - uv0) + b2 (uv2 - uv0)
 *   col = color0 + b1 (color1 - color0) + b2 (color2 - color0)
 *   out = Sample(texture of the material, U, V) * col   (RGBA, fp32)
 * Sample uses the resolve's level and per-4x4-tile filter choice (nearest
 * texel of the chosen level when minifying, else bilinear at level 0 with
 * texel centres at +0.5 and clamped edges) and Repeat wrap. Uncovered pixels
 * give 0, untextured materials sample 1. Gradients are reduced
 * deterministically (bitwise identical run to run). */
#define GLIMP_MAX_DIFF_TEXTURES 256

typedef struct glimp_diff_texture {
    uint32_t width, height, mip_levels, pad;
    uint64_t level_offset[16];         /* in texels; level m is (width>>m) x (height>>m), row-major */
} glimp_diff_texture;

typedef struct glimp_diff_desc {
    float object_to_clip[16];
    const glimp_diff_texture* textures; /* host; material texture i -> textures[i] */
    uint32_t num_textures;
    uint64_t num_texels;               /* texels holds num_texels RGBA fp32 texels */
    const float* texels;               /* device [num_texels][4] */
    const float* uv;                   /* device [scene meshlets * 64][2] */
    const float* color;                /* device [scene meshlets * 64][4] */
} glimp_diff_desc;

/* out: device [height][width][4] fp32. Asynchronous on `stream`. */
int glimp_diff_forward(glimp_context* ctx, glimp_framebuffer* fb, glimp_scene* scene, const glimp_diff_desc* desc,
                       float* out, void* stream);
/* Gradients of sum(grad_out * out): grad_texels [num_texels][4], grad_uv
 * [meshlets * 64][2], grad_color [meshlets * 64][4], all device, fully
 * overwritten. Asynchronous on `stream`; calls on one context share its
 * workspace and must not overlap. */
int glimp_diff_backward(glimp_context* ctx, glimp_framebuffer* fb, glimp_scene* scene, const glimp_diff_desc* desc,
                        const float* grad_out, float* grad_texels, float* grad_uv, float* grad_color, void* stream);

/* Several framebuffers of one scene in one launch per stage: outputs and
 * gradient inputs are the views' row-major images one after another
 * ([sum of width*height][4]); gradients are summed over all views with the
 * same deterministic reduction. desc->object_to_clip is not used. */
typedef struct glimp_diff_view {
    glimp_framebuffer* framebuffer;
    float object_to_clip[16];
} glimp_diff_view;

int glimp_diff_forward_views(glimp_context* ctx, glimp_scene* scene, const glimp_diff_desc* desc,
                             const glimp_diff_view* views, uint32_t count, float* out, void* stream);
int glimp_diff_backward_views(glimp_context* ctx, glimp_scene* scene, const glimp_diff_desc* desc,
                              const glimp_diff_view* views, uint32_t count, const float* grad_out,
                              float* grad_texels, float* grad_uv, float* grad_color, void* stream);

/* ------------------------------------------------------------ resolve */

#define GLIMP_RESOLVE_ALBEDO 0u   /* ResolveDebug(DebugLayer::BaseColor) (Shading.cpp:734-773) */
#define GLIMP_RESOLVE_LIT 1u      /* Resolve (Shading.cpp:658-727), without the skybox */
#define GLIMP_RESOLVE_NORMALS 2u             /* ResolveDebug(DebugLayer::Normals) */
#define GLIMP_RESOLVE_METALLIC_ROUGHNESS 3u  /* ResolveDebug(DebugLayer::MetallicRoughness) */
#define GLIMP_RESOLVE_MESHLET_ID 4u          /* ResolveDebug(DebugLayer::MeshletId) */
#define GLIMP_RESOLVE_TRIANGLE_ID 5u         /* ResolveDebug(DebugLayer::TriangleId) */
#define GLIMP_RESOLVE_OVERDRAW_PIXEL 6u      /* ResolveDebug(DebugLayer::OverdrawPixel); overdraw framebuffers */
#define GLIMP_RESOLVE_OVERDRAW_QUAD 7u       /* ResolveDebug(DebugLayer::OverdrawQuad); overdraw framebuffers */
#define GLIMP_MAX_LIGHTS 8u
/* glimp_resolve_desc.flags: Resolve draws its point and spot lights as
 * sprites over the image (Shading.cpp:689-727); needs world_to_clip. */
#define GLIMP_RESOLVE_LIGHT_SPRITES 0x1u

/* Light (Scene.h:52-75) with its precomputed terms. */
typedef struct glimp_light {
    uint32_t type;                     /* 0 directional, 1 point, 2 spot */
    float position[3];
    float direction[3];
    float color[3];
    float intensity;
    float inv_radius_sq, spot_scale, spot_offset;
} glimp_light;

typedef struct glimp_resolve_desc {
    uint32_t mode;                     /* GLIMP_RESOLVE_* */
    float object_to_clip[16];          /* ShadingContext::ObjectToClipMat */
    float object_to_world[9];          /* ShadingContext::ObjectToWorldMat (mat3, column-major) */
    float inv_screen_proj[16];         /* GetInverseScreenProjMatrix(WorldToClipMat, size) (Camera.h:140-146) */
    float view_pos[3];                 /* ShadingContext::ViewPos */
    float exposure;                    /* ShadingContext::Exposure */
    const glimp_light* lights;         /* host array (ShadingContext::Lights) */
    uint32_t num_lights;               /* <= GLIMP_MAX_LIGHTS */
    uint32_t flags;                    /* GLIMP_RESOLVE_LIGHT_SPRITES (lit mode) */
    float world_to_clip[16];           /* ShadingContext::WorldToClipMat, column-major (light sprites) */
    const glimp_texture* skybox;       /* ShadingContext::SkyboxTex (lit mode): an R11G11B10f octahedral map
                                        * for the pixels nothing covers (Shading.cpp:675-678), or NULL */
} glimp_resolve_desc;

/* ShadingContext::Resolve / ResolveDebug (Shading.cpp:658-773):
 * writes RGBA8 per pixel into the framebuffer's color buffer (row-major),
 * which glimp_framebuffer_color / glimp_framebuffer_read_color expose.
 * Asynchronous. */
int glimp_resolve(glimp_context* ctx, glimp_framebuffer* fb, glimp_scene* scene, const glimp_resolve_desc* desc,
                  void* stream);
/* Batches (above): adds a resolve of every view after its draw, in the same launch and the
 * same captured graph: descs holds one glimp_resolve description per view
 * of the batch, in view order; descs = NULL removes them. Light arrays are
 * copied. The replay of a frame (clear, cull, draw, resolve) is then one
 * graph launch. */
int glimp_batch_set_resolve(glimp_batch* batch, const glimp_resolve_desc* descs);
/* Device pointer to the resolved RGBA8 buffer (width*height u32, row-major). */
int glimp_framebuffer_color(glimp_framebuffer* fb, uint32_t** color);
/* Copies the resolved RGBA8 buffer to host memory (row-major). Synchronous. */
int glimp_framebuffer_read_color(glimp_framebuffer* fb, uint32_t* rgba, void* stream);

/* Perf counters (Rasterizer.h:366-379), accumulated since the last reset. */
typedef struct glimp_counters {
    uint64_t triangles_processed;      /* PerfCounter::TrianglesProcessed */
    uint64_t triangles_rasterized;     /* PerfCounter::TrianglesRasterized */
    uint64_t triangles_clipped;        /* PerfCounter::TrianglesClipped */
    uint64_t bins_rasterized;          /* PerfCounter::BinQueueFlushes */
    uint64_t fragments_covered;        /* all counters need GLIMP_DRAW_COUNTERS */
    uint64_t fragments_passed;
    uint64_t setup_records;            /* triangles that left setup */
    uint64_t bin_entries;              /* (triangle, bin) pairs */
    uint64_t dropped;                  /* queued clip triangles, clipped records or bin-overflow pairs lost
                                        * to a full queue (counted always; non-zero means the frame is
                                        * incomplete) */
    uint64_t tile_lanes;               /* 16 per 4x4 tile the bin raster evaluates (lane utilisation:
                                        * fragments_covered / tile_lanes) */
} glimp_counters;
int glimp_get_counters(glimp_context* ctx, glimp_counters* out); /* synchronous */
int glimp_reset_counters(glimp_context* ctx);

/* Kernel timing: with profiling enabled every kernel of a draw or resolve is
 * bracketed by HIP events on its stream; glimp_profile_read waits for them
 * and returns per-kernel launch counts and summed milliseconds (sorted by
 * name). The device side of the reference's DrawTime/ResolveTime counters
 * (Rasterizer.h:372-373, SWR_PERF_BEGIN/END Rasterizer.h:398-399). */
typedef struct glimp_kernel_time {
    char name[40];
    uint64_t launches;
    double total_ms;
} glimp_kernel_time;
int glimp_profile_enable(glimp_context* ctx, int enable);
int glimp_profile_read(glimp_context* ctx, glimp_kernel_time* out, uint32_t capacity, uint32_t* count, int reset);

#ifdef __cplusplus
}
#endif
#endif
