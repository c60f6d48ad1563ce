// Shared device-side structures and kernel launchers of the glimp library.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "device_math.h"
#include "profiler.h"

namespace glimp {

// Device copy of the meshlet structure-of-arrays (Scene.h:15-30).
struct DevMeshlets {
    uint32_t count;
    const float4* bounds;
    const uint8_t* num_vertices;
    const uint8_t* num_triangles;
    const uint32_t* material_id;
    const uint64_t* tangent_handedness;
    const float* positions;        // [M][3][64]
    const uint32_t* texcoords;     // [M][64]
    const uint32_t* normal_tangents;
    const uint8_t* indices;        // [M][3][128]
};

// Per material: face culling and fragment entry point as ShadeMeshlet
// derives them (Shading.cpp:302-306), plus the texture slot.
struct DevMaterial {
    int32_t texture;
    uint8_t cull_mode;             // 0 None, 1 FrontCCW
    uint8_t fs_id;                 // 0 = FS_EncodeSurfaceId<false>, 1 = <true>
    uint8_t alpha_cutoff;
    uint8_t pad;
};

struct DevTexture {                // Texture2D header (Texture.h:323-328)
    uint32_t width, height, mip_levels, num_layers;
    uint32_t row_shift, layer_stride;
    float scale_u, scale_v, scale_lerp_u, scale_lerp_v;
    int32_t mask_lerp_u, mask_lerp_v;
    uint32_t mip_offsets[16];
    const uint32_t* data;
};

// One triangle that survived setup, 32 bytes. Edge equations and the
// render box are re-derived from the snapped positions where needed.
struct TriRecord {
    uint32_t pos[3];               // snapped 1/16 px, x | y << 16 (s16 each)
    float z0, z10, z20;            // depth plane (Rasterizer.cpp:321-323)
    uint32_t rank;                 // draw order, see key layout
    uint32_t flags;                // bit 0 alpha-tested (kRecAlpha), bit 1 on the overflow list (kRecSpilled), bits 16-31 view index
};
static_assert(sizeof(TriRecord) == 32, "record size");

// What the alpha test of FS_EncodeSurfaceId<true> (Shading.cpp:309-331)
// needs of a record beyond coverage and depth, 64 bytes, at the record's own
// index in a side array that exists only for draws with alpha-tested
// materials: the 1/w terms of the perspective correction (W0 and W*/area,
// TriangleEdgeVars::Setup, Rasterizer.cpp:324-328), the vertex UVs (half2
// each), the texture and cutoff, and for clipped triangles the remap of the
// fan triangle's barycentrics onto the source triangle (TriangleClipData
// ClippedU/V, Rasterizer.cpp:461-464).
struct AlphaRecord {
    float w0, w0s, w1s, w2s;
    uint32_t uv[3];
    uint32_t tex;                  // texture index | alpha cutoff << 24
    float ru[3], rv[3];
    uint32_t pad[2];
};
static_assert(sizeof(AlphaRecord) == 64, "alpha record size");

// TriRecord::flags bit 0: the record is alpha-tested.
constexpr uint32_t kRecAlpha = 1u;
// TriRecord::flags bit 1: a bin the record covers was full, so the record is
// on the overflow list and drawn over its whole render box by the overflow
// raster; the bin lists that hold it skip it (each fragment drawn once).
constexpr uint32_t kRecSpilled = 2u;

// A triangle that needs guard-band clipping, with its clip-space vertices.
struct ClipEntry {
    uint32_t rank_base;            // rank with fan index 0
    uint32_t info;                 // outcodes | cull_mode << 8 | alpha-tested << 10
    float v[3][4];
    uint32_t view;                 // view index within the launch
    uint32_t pad;
};
static_assert(sizeof(ClipEntry) == 64, "clip entry size");

// Draw-order rank carried in the low 32 bits of a depth-test key:
//   seq[31:11] packet[10:8] clipped[7] lane[6:3] fan[2:0]
// seq = meshlets drawn into the framebuffer before this one. Within a
// meshlet this is DrawMeshletsST's order (Rasterizer.cpp:166-249): packets
// ascending, accepted lanes in lane order, then clipped sub-triangles by
// source lane and fan position.
__host__ __device__ constexpr uint32_t make_rank(uint32_t seq, uint32_t packet, uint32_t clipped,
                                                 uint32_t lane, uint32_t fan) {
    return (seq << 11) | (packet << 8) | (clipped << 7) | (lane << 3) | fan;
}

struct Planes5 {                   // 5 normalised frustum planes, xyz + w
    float p[20];
};

struct DrawEntry {                 // maps a seq range back to meshlet ids
    uint32_t seq_base;
    uint32_t count;
    uint32_t meshlet_offset;
    uint32_t pad;
};

struct Counters {                  // device counters, u64 each
    unsigned long long triangles_processed;
    unsigned long long triangles_rasterized;
    unsigned long long triangles_clipped;
    unsigned long long bins_rasterized;
    unsigned long long fragments_covered;
    unsigned long long fragments_passed;
    unsigned long long setup_records;
    unsigned long long bin_entries;
    unsigned long long dropped;    // clipped triangles, fan records or overflow pairs lost to a full queue (always counted)
    unsigned long long tile_lanes; // 16 per 4x4 tile the bin raster and the overflow raster evaluate
};

constexpr uint32_t kBinBuckets = 16;

// Per-draw scratch counters, cleared without a launch of their own:
// num_visible is zero when a draw starts (allocation clears it, the bin
// raster clears it after its last reader, a direct draw ends with a reset),
// and the cull clears the others, which only later stages fill.
struct DrawScratch {
    uint32_t num_records;          // records appended by the clipper
    uint32_t num_clip;
    uint32_t num_visible;
    uint32_t pad0;
    uint32_t num_spill;            // records on the overflow list
    uint32_t num_clip_overflow;    // meshlets with triangles that found the clip queue full
    uint32_t max_bin_count;        // the largest bin's record count (pairs offered, kept or not)
    uint32_t pad;
    uint32_t bucket[kBinBuckets];  // bins per record-count class (log2), for heavy-first raster order
};

// Occlusion inputs of CullMeshlets (Shading.cpp:775-869): the previous
// frame's depth map, a min-depth pyramid whose level m texel covers 2^(m+1)
// pixels (texutil::DownsampleDepth), and the sphere transform terms.
// depth_map == nullptr: no occlusion test.
struct Occlusion {
    const float* depth_map;        // levels 0 .. mip_levels - 2, row-major, concatenated
    uint32_t mip_levels;           // of the reference's depth texture (CreateTexture2D rule)
    uint32_t width, height;        // frame size
    uint32_t level_offset[16];     // in floats
    uint32_t level_width[16];
    float m[16];                   // objectToPrevViewMat, column-major
    float scale;                   // objectToWorldScale
    float proj[3];                 // projFactors: znear, f/aspect, -f
};

// One view of a draw or batch: which meshlets of which scene are drawn into
// which framebuffer with which matrix. A single DrawMeshlets call is a batch
// of one view; a batch renders many views with one launch per stage.
struct ViewParams {
    DevMeshlets m;
    const DevMaterial* materials;
    uint32_t num_materials;
    const DevTexture* textures;    // the scene's textures (alpha test)
    uint32_t meshlet_offset;       // ShadingContext::MeshletOffset
    uint32_t meshlet_count;
    uint32_t meshlet_base;         // first global meshlet slot of the view
    uint32_t bin_base;             // first global bin of the view
    uint32_t bins_x, bins_y;
    uint32_t seq_base;             // meshlets drawn into the framebuffer before
    int width, height;
    float guard_x, guard_y;
    int use_cull, clipping, fresh;
    float clear_depth;
    unsigned long long* keys;      // the view's vis-buffer, 4x4 tiles
    // Same layout as `keys`, all zero between draws: records that overflowed
    // a bin (and fan triangles that found no record slot) are drawn here with
    // global atomic maxima, marking the bins they reach; the raster of a
    // marked bin takes the maximum with its own keys and zeroes its pixels
    // again.
    unsigned long long* spill_keys;
    uint32_t* aux;                 // overdraw: the depth layer; deferred: colour layer, then normal/MR layer
    float o2w[9];                  // ShadingContext::ObjectToWorldMat (deferred normals), mat3 column-major
    float clip[16];
    float planes[20];
    Occlusion occ;                 // with use_cull: HiZ test of the frustum survivors
};

// Launch-wide parameters. Meshlets of all views are numbered globally
// (view v owns slots [meshlet_base, meshlet_base + meshlet_count)), and so
// are the screen bins (view v owns [bin_base, bin_base + bins_x*bins_y)).
// Bins are bin_size x bin_size pixel squares (a multiple of the 4x4 tile);
// each bin holds up to bin_capacity record indices. A record with a pair
// that finds its bin full is appended to the overflow list instead (once)
// and flagged; the overflow raster draws it over its render box into the
// view's spill_keys, marking the bins it reaches, whose rasters skip it and
// merge the spill keys. Each fragment is drawn once, and the depth-test
// maximum does not depend on the order.
struct DrawParams {
    uint32_t num_views;
    const ViewParams* views;       // device array, or null: the inline view
    ViewParams view;
    uint32_t total_meshlets;
    uint32_t total_bins;
    uint32_t flags;
    ApproxTables tables;
    uint2* visible_list;           // (global meshlet number, view) of each meshlet that passed the cull
    TriRecord* records;            // [total_meshlets * 128 + clip_record_capacity]
    AlphaRecord* alpha;            // same indexing as records; null when no view has alpha-tested materials
    uint32_t deferred;             // FS_EncodeGBuffer: every textured material runs the alpha test
    uint32_t* meshlet_counts;      // [total_meshlets] records kept per meshlet
    uint32_t clip_record_base;     // = total_meshlets * 128
    uint32_t clip_record_capacity;
    ClipEntry* clip_queue;
    uint32_t clip_capacity;
    DrawScratch* scratch;
    Counters* counters;
    uint32_t bin_size, bin_shift, bin_capacity;
    uint32_t raster_threads;       // bin raster workgroup size: 128 or 256 (512, 1024 for 32/64-pixel bins)
    uint32_t* bin_counts;          // [total_bins]
    uint32_t* bin_lists;           // [total_bins][bin_capacity]
    uint32_t* bin_order;           // [kBinBuckets][total_bins]: bins by record-count class
    uint32_t* spill;               // overflow list: record indices, one per record slot at most
    uint32_t spill_capacity;
    // [total_bins], zero between draws: set for a bin when a record of the
    // overflow list or a fan triangle of a triangle that found the clip queue
    // full was drawn into its spill keys
    uint32_t* bin_flags;
    // Triangles that find the clip queue full are kept as references in free
    // slots of their meshlet's record block and listed here as (visible slot,
    // count), [total_meshlets]; they are clipped and drawn after the binning
    // (the vis-buffer pipeline; the others count them in `dropped`).
    uint2* clip_overflow;
    uint32_t exact_overflow;
};

// Launchers (glimp_kernels.hip). `max_meshlets` bounds the meshlets of
// the largest view (grid sizing only).
// Queries the resident block counts of the looping kernels (call once per
// process on the device, outside any stream capture).
void warm_launch_shapes();
hipError_t launch_draw_direct(const DrawParams& p, hipStream_t s, Profiler* prof);
hipError_t launch_draw_binned(const DrawParams& p, hipStream_t s, Profiler* prof);
// The overdraw pipeline (FS_Overdraw): per pixel the keys hold the covered
// count in the high and the quad-helper count in the low 32 bits, and
// V.aux the depth layer.
hipError_t launch_draw_overdraw(const DrawParams& p, hipStream_t s, Profiler* prof);
hipError_t launch_decode_overdraw(const unsigned long long* counts, const uint32_t* depth, uint32_t n,
                                  uint32_t clear_color, int width, uint32_t* words, uint32_t* depth_out, int linear,
                                  hipStream_t s);
// The deferred pipeline (FS_EncodeGBuffer): the vis-buffer raster with the
// alpha test on every textured material, then the G-buffer encode of each
// pixel's winning fragment into V.aux (colour layer, then layer 2).
hipError_t launch_draw_deferred(const DrawParams& p, hipStream_t s, Profiler* prof);
// 4x4-tiled 32-bit layer -> host order (row-major when linear).
hipError_t launch_detile32(const uint32_t* src, int width, uint32_t n, uint32_t* out, int linear, hipStream_t s);
hipError_t launch_clear(unsigned long long* keys, uint32_t n, unsigned long long key, hipStream_t s);
// Zeroes a draw's scratch counters (one thread).
hipError_t launch_reset_counters(DrawScratch* scratch, hipStream_t s);
hipError_t launch_fill32(uint32_t* buf, uint32_t n, uint32_t value, hipStream_t s);
hipError_t launch_decode(const unsigned long long* keys, int width, int height, unsigned long long clear_key,
                         uint32_t clear_color, const DrawEntry* draws, uint32_t num_draws, uint32_t* ids,
                         float* depth, int linear, hipStream_t s);
hipError_t launch_cull(const DevMeshlets& m, uint32_t first, uint32_t count, const float* planes,
                       const Occlusion* occ, const ApproxTables& tables, uint16_t* bitmap, uint32_t* visible,
                       hipStream_t s);
// Min-depth pyramid of a vis-buffer into occ.depth_map (all levels).
hipError_t launch_hiz_build(const unsigned long long* keys, const Occlusion& occ, float* depth_map, hipStream_t s);

}  // namespace glimp
