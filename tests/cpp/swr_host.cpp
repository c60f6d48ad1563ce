// Host program on the C++ swr layer (include/glimp_swr.hpp): loads a scene
// written by tests/test_gpu_swr_host.py, runs the reference frame sequence
// (Main.cpp:211-252: clear, cull + draw, ResolveDebug(BaseColor)) and writes
// the surface ids, depth and resolved colour for comparison with the oracle.
//
//   swr_host <dir>
// <dir>/meta.txt: count width height tex_w tex_h tex_levels tex_layers
//                 row_shift layer_stride num_texels, then 16 mip offsets,
//                 16 object_to_clip floats, 20 plane floats
// <dir>/*.bin: the meshlet arrays and texels (see the test).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "glimp_swr.hpp"

template <class T>
static std::vector<T> load(const std::string& path, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <class T>
static void save(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: swr_host <dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    std::ifstream meta(dir + "/meta.txt");
    uint32_t count, W, H, tw, th, tl, tlayers, row_shift, layer_stride;
    size_t num_texels;
    meta >> count >> W >> H >> tw >> th >> tl >> tlayers >> row_shift >> layer_stride >> num_texels;
    uint32_t mip_offsets[16];
    for (auto& m : mip_offsets) meta >> m;
    float clip[16], planes[20];
    for (auto& c : clip) meta >> c;
    for (auto& p : planes) meta >> p;
    if (!meta) {
        std::fprintf(stderr, "bad meta.txt\n");
        return 2;
    }
    auto bounds = load<float>(dir + "/bounds.bin", (size_t)count * 4);
    auto nv = load<uint8_t>(dir + "/num_vertices.bin", count);
    auto nt = load<uint8_t>(dir + "/num_triangles.bin", count);
    auto mat = load<uint32_t>(dir + "/material_id.bin", count);
    auto th64 = load<uint64_t>(dir + "/tangent_handedness.bin", count);
    auto pos = load<float>(dir + "/positions.bin", (size_t)count * 192);
    auto tc = load<uint32_t>(dir + "/texcoords.bin", (size_t)count * 64);
    auto ntg = load<uint32_t>(dir + "/normal_tangents.bin", (size_t)count * 64);
    auto idx = load<uint8_t>(dir + "/indices.bin", (size_t)count * 384);
    auto texels = load<uint32_t>(dir + "/texels.bin", num_texels);

    try {
        using namespace glimp::swr;
        Context gpu(0);
        Rasterizer rast(gpu);
        glimp_texture_desc td{};
        td.width = tw;
        td.height = th;
        td.mip_levels = tl;
        td.num_layers = tlayers;
        td.row_shift = row_shift;
        td.layer_stride = layer_stride;
        for (int i = 0; i < 16; i++) td.mip_offsets[i] = mip_offsets[i];
        td.texels = texels.data();
        Texture2D tex(gpu, td);
        glimp_meshlets_desc md{ count, bounds.data(), nv.data(), nt.data(), mat.data(), th64.data(),
                                pos.data(), tc.data(), ntg.data(), idx.data() };
        glimp_material m0{};
        m0.texture = 0;
        m0.alpha_cutoff = 255;
        Scene scene(gpu, md, { m0 }, { &tex });

        auto fb = CreateFramebuffer(gpu, W, H);
        ShadingContext shader(scene);
        const auto& dispatch = GetDispatchTable<ShadeMeshlet, FS_EncodeSurfaceId<false>, FS_EncodeSurfaceId<true>>();
        fb.Clear(0, 0.0f);
        shader.MeshletOffset = 0;
        shader.UpdateProj(clip);
        shader.SetFrustumPlanes(planes);
        rast.DrawMeshlets(fb, count, { dispatch, &shader });

        std::vector<uint32_t> ids((size_t)W * H), depth((size_t)W * H), color((size_t)W * H);
        fb.GetPixels(0, ids.data(), W);
        fb.GetPixels(1, depth.data(), W);
        shader.ResolveDebug(rast, fb, DebugLayer::BaseColor);
        fb.GetPixels(0, color.data(), W);
        // visible count through the standalone cull entry point
        std::vector<uint16_t> bitmap((count + 15) / 16);
        uint32_t visible = shader.CullMeshlets(gpu, bitmap.data(), 0, count, planes);
        // the same frame through OverdrawShader (Main.cpp:204-213: clear to 0,
        // then ResolveDebug of an overdraw layer)
        std::vector<uint32_t> odWords((size_t)W * H), odDepth((size_t)W * H), odColor((size_t)W * H);
        fb.Clear(0, 0.0f);
        rast.DrawMeshlets(fb, count, { GetDispatchTable<ShadeMeshlet, FS_Overdraw>(), &shader });
        fb.GetPixels(0, odWords.data(), W);
        fb.GetPixels(1, odDepth.data(), W);
        shader.ResolveDebug(rast, fb, DebugLayer::OverdrawQuad);
        fb.GetPixels(0, odColor.data(), W);
        save(dir + "/od_words.bin", odWords);
        save(dir + "/od_depth.bin", odDepth);
        save(dir + "/od_color.bin", odColor);
        save(dir + "/ids.bin", ids);
        save(dir + "/depth.bin", depth);
        save(dir + "/color.bin", color);
        save(dir + "/bitmap.bin", bitmap);
        std::printf("swr_host: %ux%u, %u meshlets, %u visible\n", W, H, count, visible);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "swr_host: %s\n", e.what());
        return 1;
    }
    return 0;
}
