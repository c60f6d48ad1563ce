// glimp_swr.hpp: C++ host layer over the C ABI (glimp.h) with the shape of
// the reference renderer's swr API (src/SwRast: Rasterizer.h, Shading.h,
// Texture.h), so a host written against swr::Rasterizer / ShadingContext
// keeps its structure while the work runs on the MI355X.
//
// Header-only, C++17; link libglimp.so. Errors throw glimp::swr::Error with
// the library's message (the reference asserts instead).
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "glimp.h"

namespace glimp::swr {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc, const char* what) {
    if (rc != GLIMP_OK) throw Error(rc, std::string(what) + ": " + glimp_last_error());
}

// Owns the device context (the reference's Rasterizer owns a thread pool,
// Rasterizer.h:330-332; here the context owns device scratch).
class Context {
public:
    explicit Context(int device = 0) { check(glimp_create(&h_, device), "glimp_create"); }
    ~Context() { glimp_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    glimp_context* get() const { return h_; }

private:
    glimp_context* h_ = nullptr;
};

// Framebuffer (Rasterizer.h:10-78): the two reference layers (surface id,
// depth) are one 64-bit key per pixel on the device.
class Framebuffer {
public:
    Framebuffer(Context& ctx, uint32_t width, uint32_t height) : Width(width), Height(height) {
        check(glimp_framebuffer_create(ctx.get(), width, height, &h_), "glimp_framebuffer_create");
    }
    ~Framebuffer() { glimp_framebuffer_destroy(h_); }
    Framebuffer(const Framebuffer&) = delete;
    Framebuffer& operator=(const Framebuffer&) = delete;

    // Framebuffer::Clear (Rasterizer.h:35-38)
    void Clear(uint32_t color, float depth, void* stream = nullptr) {
        check(glimp_framebuffer_clear(h_, color, depth, stream), "glimp_framebuffer_clear");
    }
    // GetPixels (ImageHelpers.cpp:109-147): layer 0 = surface ids, overdraw
    // counters or base colour (or the resolved RGBA8 colour after a
    // resolve), layer 1 = depth bits, layer 2 = the deferred pipeline's
    // normal / metallic / roughness; `stride` in pixels.
    void GetPixels(uint32_t layer, uint32_t* dest, uint32_t stride, void* stream = nullptr) {
        std::vector<uint32_t> tmp((size_t)Width * Height);
        if (layer == 0 && resolved_) {
            check(glimp_framebuffer_read_color(h_, tmp.data(), stream), "glimp_framebuffer_read_color");
        } else {
            check(glimp_framebuffer_read_layer(h_, layer, tmp.data(), 1, stream), "glimp_framebuffer_read_layer");
        }
        for (uint32_t y = 0; y < Height; y++) std::memcpy(dest + (size_t)y * stride, &tmp[(size_t)y * Width], Width * 4);
    }
    glimp_framebuffer* get() const { return h_; }

    const uint32_t Width, Height;

private:
    friend class ShadingContext;
    friend class Rasterizer;
    glimp_framebuffer* h_ = nullptr;
    bool resolved_ = false;
};

inline Framebuffer CreateFramebuffer(Context& ctx, uint32_t w, uint32_t h) { return Framebuffer(ctx, w, h); }

// RGBA8 texture in CreateTexture2D's storage layout (Texture.h:599-636).
class Texture2D {
public:
    Texture2D(Context& ctx, const glimp_texture_desc& d) { check(glimp_texture_create(ctx.get(), &d, &h_), "glimp_texture_create"); }
    ~Texture2D() { glimp_texture_destroy(h_); }
    Texture2D(const Texture2D&) = delete;
    Texture2D& operator=(const Texture2D&) = delete;
    glimp_texture* get() const { return h_; }

private:
    glimp_texture* h_ = nullptr;
};

// The previous frame's depth map for occlusion culling (Main.cpp:54-56,
// 257-260): Build() is texutil::DownsampleDepth.
class DepthMap {
public:
    DepthMap(Context& ctx, uint32_t width, uint32_t height) {
        check(glimp_hiz_create(ctx.get(), width, height, &h_), "glimp_hiz_create");
    }
    ~DepthMap() { glimp_hiz_destroy(h_); }
    DepthMap(const DepthMap&) = delete;
    DepthMap& operator=(const DepthMap&) = delete;
    void Build(Framebuffer& fb, void* stream = nullptr) { check(glimp_hiz_build(h_, fb.get(), stream), "glimp_hiz_build"); }
    glimp_hiz* get() const { return h_; }

private:
    glimp_hiz* h_ = nullptr;
};

// Scene::Meshlets + Materials (Scene.h:92-128), resident on the device.
class Scene {
public:
    Scene(Context& ctx, const glimp_meshlets_desc& meshlets, const std::vector<glimp_material>& materials,
          const std::vector<Texture2D*>& textures) {
        check(glimp_scene_create(ctx.get(), &meshlets, &h_), "glimp_scene_create");
        std::vector<glimp_texture*> th;
        for (auto* t : textures) th.push_back(t->get());
        int rc = glimp_scene_set_materials(h_, materials.data(), (uint32_t)materials.size(), th.data(),
                                           (uint32_t)th.size());
        if (rc != GLIMP_OK) {
            glimp_scene_destroy(h_);
            check(rc, "glimp_scene_set_materials");
        }
        count_ = meshlets.count;
    }
    ~Scene() { glimp_scene_destroy(h_); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    glimp_scene* get() const { return h_; }
    uint32_t MeshletCount() const { return count_; }

private:
    glimp_scene* h_ = nullptr;
    uint32_t count_ = 0;
};

// Shader tags. In the reference, GetDispatchTable<&MeshFn, &FragFn...>
// (Rasterizer.h:342-364) instantiates the raster loop for the given
// functions; on the GPU those instantiations are compiled into the library
// as pipelines, and the tags select one.
struct ShadeMeshlet {};                            // ShadingContext::ShadeMeshlet (Shading.cpp:281-307)
template <bool AlphaTest> struct FS_EncodeSurfaceId {};  // Shading.cpp:309-331
struct FS_Overdraw {};                             // Shading.cpp:333-342
struct FS_EncodeGBuffer {};                        // Shading.cpp:344-407

struct ShaderDispatchTable {
    uint32_t pipeline;
};

template <class MeshFn, class... FragFns>
struct PipelineOf {
    static_assert(sizeof(MeshFn) == 0, "no device pipeline is built for this shader combination");
};
// ShadingContext::VisBufferShader (Shading.cpp:650-654): meshlets whose
// material has an alpha cutoff below 255 take FS_EncodeSurfaceId<true>.
template <>
struct PipelineOf<ShadeMeshlet, FS_EncodeSurfaceId<false>, FS_EncodeSurfaceId<true>> {
    static constexpr uint32_t value = GLIMP_PIPELINE_VISBUFFER;
};
// The table without the alpha-tested entry: in the reference a meshlet with
// FragmentShaderId 1 has no function in it; here it takes the alpha test.
template <>
struct PipelineOf<ShadeMeshlet, FS_EncodeSurfaceId<false>> {
    static constexpr uint32_t value = GLIMP_PIPELINE_VISBUFFER;
};

// ShadingContext::DeferredShader (Shading.cpp:655).
template <>
struct PipelineOf<ShadeMeshlet, FS_EncodeGBuffer> {
    static constexpr uint32_t value = GLIMP_PIPELINE_DEFERRED;
};
// ShadingContext::OverdrawShader (Shading.cpp:656).
template <>
struct PipelineOf<ShadeMeshlet, FS_Overdraw> {
    static constexpr uint32_t value = GLIMP_PIPELINE_OVERDRAW;
};

template <class MeshFn, class... FragFns>
const ShaderDispatchTable& GetDispatchTable() {
    static const ShaderDispatchTable table{ PipelineOf<MeshFn, FragFns...>::value };
    return table;
}

// DebugLayer (Shading.h:8); the overdraw layers read a framebuffer drawn
// with the overdraw table, the others a vis-buffer.
enum class DebugLayer { None, BaseColor, Normals, MetallicRoughness, MeshletId, TriangleId, OverdrawPixel, OverdrawQuad };

struct Lighting {
    float object_to_world[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
    float inv_screen_proj[16] = {};
    float world_to_clip[16] = {};   // ShadingContext::WorldToClipMat (light sprites)
    bool light_sprites = true;      // Resolve draws point and spot lights (Shading.cpp:689-727)
    float view_pos[3] = {};
    float exposure = 1.0f;
    std::vector<glimp_light> lights;
    const Texture2D* SkyboxTex = nullptr;   // ShadingContext::SkyboxTex: an R11G11B10f octahedral map
};

class Rasterizer;

// ShadingContext (Shading.h:10-60): per-draw matrices and the scene the
// shaders read.
class ShadingContext {
public:
    explicit ShadingContext(Scene& scene) : scene_(&scene) {}

    float ObjectToClipMat[16] = {};
    float ObjectToWorldMat[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };   // ResolveDebug normals
    uint32_t MeshletOffset = 0;

    // UpdateProj (Shading.h:35-39) with the product already formed.
    void UpdateProj(const float objectToClip[16]) { std::memcpy(ObjectToClipMat, objectToClip, sizeof(ObjectToClipMat)); }
    // CullMeshlets runs inside the next draw with these planes (five
    // normalised planes, Shading.cpp:783-791); nullptr disables culling.
    void SetFrustumPlanes(const float* planes) {
        hasPlanes_ = planes != nullptr;
        if (planes) std::memcpy(planes_, planes, sizeof(planes_));
    }
    // ShadingContext::CullMeshlets frustum branch (Shading.cpp:775-869).
    uint32_t CullMeshlets(Context& ctx, uint16_t* bitmap, uint32_t first, uint32_t count, const float planes[20],
                          void* stream = nullptr) {
        uint32_t visible = 0;
        check(glimp_cull_meshlets(ctx.get(), scene_->get(), first, count, planes, bitmap, &visible, stream),
              "glimp_cull_meshlets");
        return visible;
    }
    // CullMeshlets with the occlusion branch (Shading.cpp:811-840): the terms
    // the reference derives from projMat, modelMat and prevViewMat are passed
    // in `occ` (see glimp_occlusion).
    uint32_t CullMeshlets(Context& ctx, uint16_t* bitmap, uint32_t first, uint32_t count, const float planes[20],
                          const glimp_occlusion& occ, void* stream = nullptr) {
        uint32_t visible = 0;
        check(glimp_cull_meshlets_occlusion(ctx.get(), scene_->get(), first, count, planes, &occ, bitmap, &visible,
                                            stream),
              "glimp_cull_meshlets_occlusion");
        return visible;
    }
    // Resolve (Shading.cpp:658-732).
    void Resolve(Rasterizer& rast, Framebuffer& fb, const Lighting& L, void* stream = nullptr);
    // ResolveDebug (Shading.cpp:734-773).
    void ResolveDebug(Rasterizer& rast, Framebuffer& fb, DebugLayer layer, void* stream = nullptr);

    Scene* GetScene() const { return scene_; }
    const float* Planes() const { return hasPlanes_ ? planes_ : nullptr; }

private:
    Scene* scene_;
    float planes_[20] = {};
    bool hasPlanes_ = false;
};

struct ShaderBinding {
    const ShaderDispatchTable& Dispatch;
    ShadingContext* Context;
};

// Rasterizer (Rasterizer.h:200-240).
class Rasterizer {
public:
    explicit Rasterizer(Context& ctx) : ctx_(&ctx) {}

    bool EnableClipping = true;     // Rasterizer.h:207
    bool EnableGuardband = true;    // Rasterizer.h:208
    uint32_t BinSize = 0;           // 0 = library default
    uint32_t BinCapacity = 0;

    // Rasterizer::DrawMeshlets (Rasterizer.h:213).
    void DrawMeshlets(Framebuffer& fb, uint32_t count, ShaderBinding shader, void* stream = nullptr) {
        ShadingContext& sc = *shader.Context;
        glimp_draw_desc d;
        std::memset(&d, 0, sizeof(d));
        d.meshlet_offset = sc.MeshletOffset;
        d.meshlet_count = count;
        std::memcpy(d.object_to_clip, sc.ObjectToClipMat, sizeof(d.object_to_clip));
        d.frustum_planes = sc.Planes();
        d.flags = (EnableClipping ? 0u : GLIMP_DRAW_NO_CLIPPING) | (EnableGuardband ? 0u : GLIMP_DRAW_NO_GUARDBAND);
        d.pipeline = shader.Dispatch.pipeline;
        std::memcpy(d.object_to_world, sc.ObjectToWorldMat, sizeof(d.object_to_world));
        d.bin_size = BinSize;
        d.bin_capacity = BinCapacity;
        check(glimp_draw_meshlets(ctx_->get(), fb.get(), sc.GetScene()->get(), &d, stream), "glimp_draw_meshlets");
        fb.resolved_ = false;
    }
    Context& GetContext() const { return *ctx_; }

private:
    Context* ctx_;
};

inline void ShadingContext::Resolve(Rasterizer& rast, Framebuffer& fb, const Lighting& L, void* stream) {
    glimp_resolve_desc d;
    std::memset(&d, 0, sizeof(d));
    d.mode = GLIMP_RESOLVE_LIT;
    std::memcpy(d.object_to_clip, ObjectToClipMat, sizeof(d.object_to_clip));
    std::memcpy(d.object_to_world, L.object_to_world, sizeof(d.object_to_world));
    std::memcpy(d.inv_screen_proj, L.inv_screen_proj, sizeof(d.inv_screen_proj));
    std::memcpy(d.view_pos, L.view_pos, sizeof(d.view_pos));
    d.exposure = L.exposure;
    d.lights = L.lights.data();
    d.num_lights = (uint32_t)L.lights.size();
    d.flags = L.light_sprites ? GLIMP_RESOLVE_LIGHT_SPRITES : 0u;
    d.skybox = L.SkyboxTex ? L.SkyboxTex->get() : nullptr;
    std::memcpy(d.world_to_clip, L.world_to_clip, sizeof(d.world_to_clip));
    check(glimp_resolve(rast.GetContext().get(), fb.get(), scene_->get(), &d, stream), "glimp_resolve");
    fb.resolved_ = true;
}

inline void ShadingContext::ResolveDebug(Rasterizer& rast, Framebuffer& fb, DebugLayer layer, void* stream) {
    glimp_resolve_desc d;
    std::memset(&d, 0, sizeof(d));
    switch (layer) {
    case DebugLayer::BaseColor: d.mode = GLIMP_RESOLVE_ALBEDO; break;
    case DebugLayer::Normals: d.mode = GLIMP_RESOLVE_NORMALS; break;
    case DebugLayer::MetallicRoughness: d.mode = GLIMP_RESOLVE_METALLIC_ROUGHNESS; break;
    case DebugLayer::MeshletId: d.mode = GLIMP_RESOLVE_MESHLET_ID; break;
    case DebugLayer::TriangleId: d.mode = GLIMP_RESOLVE_TRIANGLE_ID; break;
    case DebugLayer::OverdrawPixel: d.mode = GLIMP_RESOLVE_OVERDRAW_PIXEL; break;
    case DebugLayer::OverdrawQuad: d.mode = GLIMP_RESOLVE_OVERDRAW_QUAD; break;
    default: throw Error(GLIMP_E_UNSUPPORTED, "ResolveDebug: no such layer");
    }
    std::memcpy(d.object_to_clip, ObjectToClipMat, sizeof(d.object_to_clip));
    std::memcpy(d.object_to_world, ObjectToWorldMat, sizeof(d.object_to_world));
    d.exposure = 1.0f;
    check(glimp_resolve(rast.GetContext().get(), fb.get(), scene_->get(), &d, stream), "glimp_resolve");
    fb.resolved_ = true;
}

}  // namespace glimp::swr
