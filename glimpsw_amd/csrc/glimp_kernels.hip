// Meshlet raster kernels for gfx950 (wave64).
//
// One launch per stage covers every view of a draw or batch (a single
// DrawMeshlets call is a batch of one view, Rasterizer::DrawMeshlets,
// Rasterizer.h:213):
//   k_begin       zero the per-launch counters and bin counts
//   k_cull        frustum test of bound spheres, one lane per meshlet
//   k_mesh_setup  one wave per meshlet: vertex transform into LDS, outcodes,
//                 triangle setup; survivors become 32-byte TriRecords in the
//                 meshlet's own slots, guard-band cases go to a clip queue
//   k_clip        one wave per queued triangle: wave-parallel
//                 Sutherland-Hodgman, fan triangulation, setup
//   k_bin_scatter (record, screen bin) pairs into fixed-capacity bins; a
//                 record with a pair that finds its bin full goes to the
//                 overflow list
//   k_bin_order   bins ordered by record count, most first; extra blocks
//                 draw the overflow list into the views' spill keys
//   k_bin_raster  one workgroup per bin: LDS-resident 64-bit keys from the
//                 bin's list, merged with its spill keys when marked
//   k_raster_direct  without binning, all records: global 64-bit atomicMax
//
// Depth test. The reference keeps the nearest fragment (reverse-Z, strictly
// greater depth wins, Shading.cpp:313) and, with one thread, resolves ties in
// draw order. Each fragment here carries key = depth_bits << 32 | ~rank,
// rank being its position in that draw order; all depths that can pass are
// positive, so the largest key is the reference's surviving fragment and the
// result does not depend on the order the GPU produces fragments in.
//
// Alpha test. Meshlets whose material has an alpha cutoff below 255 run
// FS_EncodeSurfaceId<true> (Shading.cpp:309-331): setup writes a 64-byte
// AlphaRecord beside each of their records, and the raster kernels' Alpha
// variants drop fragments whose texel fails the cutoff. Which fragments pass
// is a function of the triangle and its 4x4 tile alone, so the key maximum
// still reproduces the reference's sequential result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "internal.h"
#include "normals.h"
#include "sampler.h"

namespace glimp {

namespace {

constexpr int kWave = 64;
constexpr uint32_t kPersistentBlocks = 2048; // looping grids: at most 8 blocks of 256 threads per CU

// Blocks of a looping kernel that are resident at once on this device (the
// grid of a kernel whose waves loop over a work list; more blocks would only
// queue behind the first ones and finish last).
template <auto Kernel, uint32_t Threads>
uint32_t resident_blocks() {
    static uint32_t cached = 0;     // one per kernel (devices of one kind assumed)
    if (!cached) {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, (int)Threads, 0) != hipSuccess ||
            per_cu <= 0 || cus <= 0)
            return kPersistentBlocks;
        cached = std::min<uint32_t>((uint32_t)(per_cu * cus), kPersistentBlocks);
    }
    return cached;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive scan across the wave with DPP moves (no LDS round trips):
// row_shr 1, 2, 4, 8 within each row of 16 lanes, then row_bcast:15 and
// row_bcast:31 carry the row totals into the rows above. Lanes a move has no
// source for, and rows its row mask leaves out, take `identity`.
template <class Op>
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t identity, Op op) {
    const int id = (int)identity;
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x111, 0xF, 0xF, false));   // row_shr:1
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x112, 0xF, 0xF, false));   // row_shr:2
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x114, 0xF, 0xF, false));   // row_shr:4
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x118, 0xF, 0xF, false));   // row_shr:8
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x142, 0xA, 0xF, false));   // row_bcast:15
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x143, 0xC, 0xF, false));   // row_bcast:31
    return v;
}

__device__ __forceinline__ uint32_t wave_scan_add(uint32_t v) {
    return wave_scan(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

// Wave-uniform minimum / maximum (lane 63 of the inclusive scan).
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
    v = wave_scan(v, 0xFFFFFFFFu, [](uint32_t a, uint32_t b) { return min(a, b); });
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    v = wave_scan(v, 0u, [](uint32_t a, uint32_t b) { return max(a, b); });
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Render bounding box of a snapped triangle (GetBoundingBox and
// GetRenderBoundingBox, Rasterizer.cpp:331-351), including the carry of the
// 32-bit add across the packed halves. Returns min/max as packed x | y << 16.
__device__ __forceinline__ void render_bbox(uint32_t p0, uint32_t p1, uint32_t p2, int hw, int hh,
                                            uint32_t& bmin, uint32_t& bmax) {
    int32_t mnx = min(min(lo16(p0), lo16(p1)), lo16(p2));
    int32_t mny = min(min(hi16(p0), hi16(p1)), hi16(p2));
    int32_t mxx = max(max(lo16(p0), lo16(p1)), lo16(p2));
    int32_t mxy = max(max(hi16(p0), hi16(p1)), hi16(p2));
    uint32_t mn = pk16(mnx, mny) + 0x00070007u;
    uint32_t mx = pk16(mxx, mxy) + 0x00070007u;
    int32_t ax = lo16(mn) >> 4, ay = hi16(mn) >> 4;
    int32_t bx = lo16(mx) >> 4, by = hi16(mx) >> 4;
    int W = hw * 2, H = hh * 2;
    ax = min(max((int32_t)(int16_t)(ax + hw), 0), W);
    ay = min(max((int32_t)(int16_t)(ay + hh), 0), H);
    bx = min(max((int32_t)(int16_t)(bx + hw), 0), W);
    by = min(max((int32_t)(int16_t)(by + hh), 0), H);
    bmin = pk16(ax, ay) & ~0x00030003u;
    bmax = (pk16(bx, by) + 0x00030003u) & ~0x00030003u;
}

// TrianglePacket::Setup (Rasterizer.cpp:256-289) for one triangle plus the
// float part of TriangleEdgeVars::Setup (Rasterizer.cpp:296-329). Fills the
// record's positions and depth plane and `wt` with the perspective terms
// {W0, W0*s, W1*s, W2*s}, s = 16/area (Rasterizer.cpp:324-328); returns
// whether the triangle is kept.
__device__ bool setup_triangle(const ApproxTables& T, const float v[3][4], int hw, int hh, int cull_mode,
                               TriRecord& rec, float (&wt)[4]) {
    float rw[3], x[3], y[3], z[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        rw[k] = frcp(T, v[k][3]);
        x[k] = v[k][0] * rw[k];
        y[k] = v[k][1] * rw[k];
        z[k] = v[k][2] * rw[k];
    }
    float det = __builtin_fmaf(x[2] - x[0], y[1] - y[0], (x[0] - x[1]) * (y[2] - y[0]));
    if (cull_mode != 1) {
        bool flip = (cull_mode == 2) ? true : (det < 0.0f);
        det = flip ? -det : det;
    }
    if (!(det > 0.0f)) return false;
    float fsx = (float)(hw * 16), fsy = (float)(hh * 16);
    int32_t X[3], Y[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        X[k] = cvt_rne(x[k] * fsx);
        Y[k] = cvt_rne(y[k] * fsy);
        rec.pos[k] = ((uint32_t)X[k] & 0xFFFFu) | ((uint32_t)Y[k] << 16);
    }
    uint32_t bmin, bmax;
    render_bbox(rec.pos[0], rec.pos[1], rec.pos[2], hw, hh, bmin, bmax);
    if (lo16(bmin) >= lo16(bmax) || hi16(bmin) >= hi16(bmax)) return false;
    // integer doubled area of the snapped triangle, made positive
    int32_t x0 = lo16(rec.pos[0]), y0 = hi16(rec.pos[0]);
    int32_t x1 = lo16(rec.pos[1]), y1 = hi16(rec.pos[1]);
    int32_t x2 = lo16(rec.pos[2]), y2 = hi16(rec.pos[2]);
    uint32_t A01 = (uint32_t)(y1 - y0), B01 = (uint32_t)(x0 - x1);
    uint32_t A20 = (uint32_t)(y0 - y2), B20 = (uint32_t)(x2 - x0);
    int32_t idet = (int32_t)(B20 * A01 - B01 * A20);
    if (idet < 0) idet = (int32_t)(0u - (uint32_t)idet);
    float rcpArea = fdiv(T, 16.0f, (float)idet);
    rec.z0 = z[0];
    rec.z10 = (z[1] - z[0]) * rcpArea;
    rec.z20 = (z[2] - z[0]) * rcpArea;
    wt[0] = rw[0];
    wt[1] = rw[0] * rcpArea;
    wt[2] = rw[1] * rcpArea;
    wt[3] = rw[2] * rcpArea;
    return true;
}

// Edge equations of a record (TriangleEdgeVars::Setup, Rasterizer.cpp:291-318).
struct Edges {
    int32_t e0, e1, e2;            // value at pixel (0,0)
    int32_t a12, a20, a01;         // x steps
    int32_t b12, b20, b01;         // y steps
};

__device__ __forceinline__ int32_t compute_edge(int32_t a, int32_t x, int32_t b, int32_t y) {
    uint32_t w = (uint32_t)a * (uint32_t)x + (uint32_t)b * (uint32_t)y;
    w += (a > 0 || (a == 0 && b > 0)) ? 0u : 0xFFFFFFFFu;
    return (int32_t)w >> 4;
}

__device__ __forceinline__ Edges edge_setup(uint32_t p0, uint32_t p1, uint32_t p2, int hw, int hh) {
    int32_t x0 = lo16(p0), y0 = hi16(p0);
    int32_t x1 = lo16(p1), y1 = hi16(p1);
    int32_t x2 = lo16(p2), y2 = hi16(p2);
    int32_t A01 = y1 - y0, B01 = x0 - x1;
    int32_t A12 = y2 - y1, B12 = x1 - x2;
    int32_t A20 = y0 - y2, B20 = x2 - x0;
    int32_t det = (int32_t)((uint32_t)B20 * (uint32_t)A01 - (uint32_t)B01 * (uint32_t)A20);
    if (det < 0) {
        A01 = -A01; B01 = -B01; A12 = -A12; B12 = -B12; A20 = -A20; B20 = -B20;
    }
    int32_t sx = -(hw << 4) + 8, sy = -(hh << 4) + 8;
    Edges e;
    e.e0 = compute_edge(A12, sx - x1, B12, sy - y1);
    e.e1 = compute_edge(A20, sx - x2, B20, sy - y2);
    e.e2 = compute_edge(A01, sx - x0, B01, sy - y0);
    e.a12 = A12; e.a20 = A20; e.a01 = A01;
    e.b12 = B12; e.b20 = B20; e.b01 = B01;
    return e;
}

__device__ __forceinline__ uint32_t pixel_offset(uint32_t W, uint32_t x, uint32_t y) {
    return ((x & ~3u) << 2) + (y & ~3u) * W + (x & 3u) + ((y & 3u) << 2);
}

// ---------------------------------------------------------------- views

// The view table is constant while a launch runs: reading it through the
// constant address space lets a wave-uniform view index become scalar loads
// (waited on apart from the vector loads in flight, which a prefetch would
// otherwise have to finish first).
__device__ __forceinline__ const ViewParams& view_of(const DrawParams& p, uint32_t v) {
    if (!p.views) return p.view;
    const __attribute__((address_space(4))) ViewParams* t = (const __attribute__((address_space(4))) ViewParams*)p.views;
    return *(const ViewParams*)(t + v);
}

// View owning global meshlet slot g (views are ordered by meshlet_base).
__device__ __forceinline__ uint32_t view_of_meshlet(const DrawParams& p, uint32_t g) {
    uint32_t lo = 0, hi = p.num_views;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (view_of(p, mid).meshlet_base <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// View owning global bin b.
__device__ __forceinline__ uint32_t view_of_bin(const DrawParams& p, uint32_t b) {
    uint32_t lo = 0, hi = p.num_views;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (view_of(p, mid).bin_base <= b) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool counting(const DrawParams& p) { return (p.flags & 8u) != 0; }

// ---------------------------------------------------------------- cull

// ProjectSphere (Shading.cpp:264-279) and the occlusion test of
// CullMeshlets (Shading.cpp:811-840) for one bound sphere that passed the
// frustum test, in the operation order of the reference's optimised build:
// the sphere's screen box picks the depth map level whose texels are about
// half the box, and the meshlet stays visible when the sphere's nearest
// depth lies in front of the farthest depth under the box. Spheres reaching
// the near plane are kept.
__device__ __forceinline__ bool hiz_visible(const ApproxTables& T, const Occlusion& o, float4 b) {
    const float* m = o.m;
    const float cx = __builtin_fmaf(b.x, m[0], __builtin_fmaf(b.y, m[4], __builtin_fmaf(b.z, m[8], m[12])));
    const float cy = __builtin_fmaf(b.x, m[1], __builtin_fmaf(b.y, m[5], __builtin_fmaf(b.z, m[9], m[13])));
    const float cz = __builtin_fmaf(b.x, m[2], __builtin_fmaf(b.y, m[6], __builtin_fmaf(b.z, m[10], m[14])));
    const float r = b.w * o.scale;
    const float nz = -cz;
    if (!((r + o.proj[0]) <= nz)) return true;
    const float crx = cx * r, cry = cy * r, crz = r * nz;
    const float ax = __builtin_fmaf(cx, cx, -(r * r)) + cz * cz, ay = __builtin_fmaf(cy, cy, -(r * r)) + cz * cz;
    const float vx = ax * rsqrt14(T, ax), vy = ay * rsqrt14(T, ay);
    const float sxh = o.proj[1] * 0.5f, syh = o.proj[2] * 0.5f;
    const float b0 = __builtin_fmaf((vx * cx - crz) * rcp14(T, vx * nz + crx), sxh, 0.5f);
    const float b1 = __builtin_fmaf((vy * cy + crz) * rcp14(T, vy * nz - cry), syh, 0.5f);
    const float b2 = __builtin_fmaf((vx * cx + crz) * rcp14(T, vx * nz - crx), sxh, 0.5f);
    const float b3 = __builtin_fmaf((vy * cy - crz) * rcp14(T, vy * nz + cry), syh, 0.5f);
    const float fw = (float)o.width, fh = (float)o.height;
    const float sx = (b2 - b0) * fw, sy = (b3 - b1) * fh;
    const float mx = (sx > sy) ? sx : sy;
    int32_t lv = (int32_t)(fbits(mx) - (127u << 23)) >> 23;
    lv = min(max(lv, 1), (int32_t)o.mip_levels - 1);
    const int32_t x0 = max(cvt_trunc(b0 * fw), 0) >> lv;
    const int32_t y0 = max(cvt_trunc(b1 * fh), 0) >> lv;
    const int32_t x1 = min(cvt_trunc(b2 * fw), (int32_t)o.width - 1) >> lv;
    const int32_t y1 = min(cvt_trunc(b3 * fh), (int32_t)o.height - 1) >> lv;
    const float dsphere = fdiv(T, o.proj[0], -(cz + r));
    // level lv - 1 of the depth map covers 2^lv pixels per texel; the range
    // lies inside it (x1 <= (width - 1) >> lv, x0 >= 0)
    const float* D = o.depth_map + o.level_offset[lv - 1];
    const int32_t w = (int32_t)o.level_width[lv - 1];
    float vis = 3.40282347e38f;
    for (int32_t y = y0; y <= y1; y++)
        for (int32_t x = x0; x <= x1; x++) vis = fminf(vis, D[y * w + x]);
    return dsphere > vis;
}

// Frustum test of bound spheres (Shading.cpp:797-809) over the global
// meshlet numbering. Survivors are compacted into visible_list so later
// stages launch work only for them: a block covers kCullPerBlock consecutive
// meshlets and claims its slice of the list with one atomic.
// PerThread meshlets per thread: 1 for small draws (more blocks in flight),
// 4 for large batches (fewer atomics on the survivor count).
template <uint32_t PerThread>
__global__ void __launch_bounds__(256) k_cull(DrawParams p) {
    constexpr uint32_t kCullPerThread = PerThread;
    constexpr uint32_t kCullPerBlock = 256 * kCullPerThread;
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base, s_view0;
    const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
    const uint32_t first = blockIdx.x * kCullPerBlock;
    // counters the stages after the cull fill; the previous draw's last
    // readers are done (num_visible is cleared after its last reader)
    if (blockIdx.x == 0 && threadIdx.x < kBinBuckets) p.scratch->bucket[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p.scratch->num_clip = 0;
        p.scratch->num_records = 0;
        p.scratch->num_spill = 0;
        p.scratch->num_clip_overflow = 0;
        p.scratch->max_bin_count = 0;
    }
    // the block's meshlets are consecutive: one search for the first view,
    // then each thread walks forward from it
    if (threadIdx.x == 0) s_view0 = view_of_meshlet(p, first);
    __syncthreads();
    uint32_t v = s_view0;
    bool vis[kCullPerThread];
    uint32_t views[kCullPerThread];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCullPerThread; k++) {
        const uint32_t g = first + k * 256 + threadIdx.x;
        bool ok = false;
        if (g < p.total_meshlets) {
            while (v + 1 < p.num_views && view_of(p, v + 1).meshlet_base <= g) v++;
            const ViewParams& V = view_of(p, v);
            ok = true;
            if (V.use_cull) {
                float4 b = V.m.bounds[V.meshlet_offset + (g - V.meshlet_base)];
#pragma unroll
                for (int j = 0; j < 5; j++) {
                    const float* P = V.planes + 4 * j;
                    float d = __builtin_fmaf(b.x, P[0], __builtin_fmaf(b.y, P[1], b.z * P[2])) + P[3];
                    ok = ok && (d > -b.w);
                }
                if (ok && V.occ.depth_map) ok = hiz_visible(p.tables, V.occ, b);
            }
        }
        vis[k] = ok;
        views[k] = v;
        mine += ok ? 1u : 0u;
    }
    const uint32_t incl = wave_scan_add(mine);
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = total ? atomicAdd(&p.scratch->num_visible, total) : 0;
    }
    __syncthreads();
    uint32_t pos = s_base + incl - mine;
    for (uint32_t w = 0; w < wid; w++) pos += s_wave[w];
#pragma unroll
    for (uint32_t k = 0; k < kCullPerThread; k++)
        if (vis[k]) p.visible_list[pos++] = make_uint2(first + k * 256 + threadIdx.x, views[k]);
}

// Standalone cull of one meshlet range (ShadingContext::CullMeshlets).
__global__ void k_cull_range(DevMeshlets m, uint32_t first, uint32_t count, Planes5 pl, Occlusion occ,
                             ApproxTables tables, uint16_t* bitmap, uint32_t* visible) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool vis = false;
    if (i < count) {
        float4 b = m.bounds[first + i];
        vis = true;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const float* P = pl.p + 4 * k;
            float d = __builtin_fmaf(b.x, P[0], __builtin_fmaf(b.y, P[1], b.z * P[2])) + P[3];
            vis = vis && (d > -b.w);
        }
        if (vis && occ.depth_map) vis = hiz_visible(tables, occ, b);
    }
    unsigned long long mask = __ballot(vis);
    uint32_t lane = lane_id();
    if ((lane & 15) == 0 && i < count) bitmap[i / 16] = (uint16_t)(mask >> (lane & 48));
    if (visible && lane == 0) {
        uint32_t c = __popcll(mask);
        if (c) atomicAdd(visible, c);
    }
}

// texutil::DownsampleDepth (ImageHelpers.cpp:149-247), one level per launch:
// level 0 texel = minimum depth of its 2x2 pixels of the vis-buffer, level
// lv texel = minimum of its 2x2 texels of level lv - 1; pixels and texels
// outside the frame count as FLT_MAX. Thread = texel.
__global__ void __launch_bounds__(256) k_hiz_level(const unsigned long long* keys, uint32_t W, uint32_t H,
                                                   uint32_t lv, const float* prev, uint32_t prev_w, float* out,
                                                   uint32_t w, uint32_t h) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w * h) return;
    const uint32_t j = t / w, i = t - j * w;
    float v = 3.40282347e38f;
    if (lv == 0) {
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t x = 2 * i + (k & 1), y = 2 * j + (k >> 1);
            if (x < W && y < H) v = fminf(v, bitsf((uint32_t)(keys[pixel_offset(W, x, y)] >> 32)));
        }
    } else {
        const uint32_t ph = ((H - 1) >> lv) + 1;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t x = 2 * i + (k & 1), y = 2 * j + (k >> 1);
            if (x < prev_w && y < ph) v = fminf(v, prev[y * prev_w + x]);
        }
    }
    out[j * w + i] = v;
}

// ---------------------------------------------------------------- mesh + setup

// Record layout: the i-th visible meshlet owns records [i*128, i*128 +
// meshlet_counts[i]); clipped sub-triangles are appended after all meshlet
// slots. Setup needs no global counter: a wave compacts its survivors with a
// ballot and writes them into its own slots. Waves loop over the visible
// list, so the grid is sized to the machine, not to the scene.
// Inputs of one meshlet for a wave (lane = vertex), loaded one meshlet
// ahead of its setup so the loads overlap the previous meshlet's work.
struct MeshletFetch {
    uint32_t g, v, mi, nt, mat;
    float px, py, pz;
    uint32_t idx[2];               // this lane's triangle of each round: v0 | v1 << 8 | v2 << 16
};

// Device-memory pointers reached through a view (whose table may be a
// kernel argument or device memory, so the compiler cannot tell): loads
// through them as global loads, not flat ones (a flat load also counts
// against the LDS wait counter, so every LDS wait would wait for it too).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* ptr) {
    return (const __attribute__((address_space(1))) T*)ptr;
}

__device__ __forceinline__ MeshletFetch fetch_meshlet(const DrawParams& p, uint32_t i, uint32_t lane) {
    MeshletFetch f;
    const uint2 gv = p.visible_list[i];
    // one meshlet per wave: its number and view are wave-uniform (scalar)
    f.g = (uint32_t)__builtin_amdgcn_readfirstlane((int)gv.x);
    f.v = (uint32_t)__builtin_amdgcn_readfirstlane((int)gv.y);
    const ViewParams& V = view_of(p, f.v);
    f.mi = V.meshlet_offset + (f.g - V.meshlet_base);
    f.nt = as_global(V.m.num_triangles)[f.mi];
    f.mat = as_global(V.m.material_id)[f.mi];
    const auto P = as_global(V.m.positions + (size_t)f.mi * 192);
    f.px = P[lane];
    f.py = P[64 + lane];
    f.pz = P[128 + lane];
    const auto I = as_global(V.m.indices + (size_t)f.mi * 384);
#pragma unroll
    for (int r = 0; r < 2; r++)
        f.idx[r] = (uint32_t)I[r * 64 + lane] | ((uint32_t)I[128 + r * 64 + lane] << 8) |
                   ((uint32_t)I[256 + r * 64 + lane] << 16);
    return f;
}

// A triangle to clip, wave-uniform: the clip queue's entry, or the values
// setup broadcasts for a triangle that found the queue full.
struct ClipSrc {
    uint32_t rank_base, info, view;
    float v[3][4];
};

__device__ __forceinline__ ClipSrc clip_src(const ClipEntry& ce) {
    ClipSrc c;
    c.rank_base = ce.rank_base;
    c.info = ce.info;
    c.view = ce.view;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 4; j++) c.v[k][j] = ce.v[k][j];
    return c;
}

// A triangle that found the clip queue full, kept in a free record slot of
// its own meshlet (slot 127 - j for the meshlet's j-th such triangle: kept
// records and clipped triangles together never exceed the meshlet's 128
// slots). Its clip-space vertices are recomputed from the meshlet.
struct ClipRef {
    uint32_t rank_base;
    uint32_t info;                 // as ClipEntry::info
    uint32_t meshlet;              // meshlet index in the scene
    uint32_t tri;                  // triangle index in the meshlet
    uint32_t view;
    uint32_t pad[3];
};
static_assert(sizeof(ClipRef) == sizeof(TriRecord), "a clip reference takes one record slot");

// `Alpha`: the launch has alpha-tested materials (p.alpha is set).
template <bool Alpha>
__device__ __forceinline__ void setup_meshlet(const DrawParams& p, const ApproxTables& AT, uint32_t i, uint32_t lane,
                                              float4 (&lds)[64], const MeshletFetch& f) {
    const uint32_t g = f.g, v = f.v;           // global meshlet number, view
    const ViewParams& V = view_of(p, v);
    const uint32_t w = g - V.meshlet_base;     // meshlet index within the view
    const int hw = V.width / 2, hh = V.height / 2;

    const uint32_t nt = f.nt;
    int cull_mode = 1;
    bool alpha = false;            // FragmentShaderId 1 with a texture to test against
    uint32_t alpha_tex = 0;
    {
        const uint32_t mat = f.mat;
        if (mat != 0xFFFFFFFFu && mat < V.num_materials) {
            const DevMaterial dm = V.materials[mat];
            cull_mode = dm.cull_mode;
            // FS_EncodeSurfaceId<true> for FragmentShaderId 1; FS_EncodeGBuffer
            // tests every textured material (Shading.cpp:357-366)
            alpha = Alpha && (dm.fs_id == 1 || p.deferred) && dm.texture >= 0;
            alpha_tex = (uint32_t)dm.texture | ((uint32_t)dm.alpha_cutoff << 24);
        }
        // ShadeMeshlet (Shading.cpp:295-298): lane = vertex, its clip-space
        // position one 16-byte LDS entry (a triangle gathers three)
        __builtin_amdgcn_wave_barrier();   // the previous meshlet's reads are done
        lds[lane] = make_float4(xform(V.clip, 0, f.px, f.py, f.pz), xform(V.clip, 1, f.px, f.py, f.pz),
                                xform(V.clip, 2, f.px, f.py, f.pz), xform(V.clip, 3, f.px, f.py, f.pz));
        __builtin_amdgcn_wave_barrier();
    }
    const uint32_t seq = V.seq_base + w;
    TriRecord* out = p.records + (size_t)i * 128;
    const uint32_t flags_hi = v << 16;
    uint32_t kept = 0, clipped = 0, overflowed = 0;

    for (uint32_t round = 0; round < 2; round++) {
        if (round * 64 >= nt) break;
        const uint32_t t = round * 64 + lane;
        bool valid = t < nt;
        bool accept = false, nontrivial = false;
        uint32_t partial = 0;
        float vv[3][4];
        if (valid) {
            uint8_t combined = 255;
            bool trivial = true;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                uint32_t ix = (f.idx[round] >> (8 * k)) & 255u;
                uint32_t vi = ix < 32 ? ix : 32 + (ix & 31);   // gather_preload64 (SIMD.h:219-227)
                {
                    const float4 q = lds[vi];
                    vv[k][0] = q.x; vv[k][1] = q.y; vv[k][2] = q.z; vv[k][3] = q.w;
                }
                float x = vv[k][0], y = vv[k][1], z = vv[k][2], ww = vv[k][3];
                uint32_t oc = 0;
                oc |= (x < -ww) ? 1u : 0u;
                oc |= (x > ww) ? 2u : 0u;
                oc |= (y < -ww) ? 4u : 0u;
                oc |= (y > ww) ? 8u : 0u;
                oc |= (z < -ww) ? 16u : 0u;
                oc |= (z > ww) ? 32u : 0u;
                partial |= oc;
                combined &= (uint8_t)oc;
                trivial = trivial && (fabsf(x) < ww * V.guard_x) && (fabsf(y) < ww * V.guard_y);
            }
            trivial = trivial && ((partial & 0x30u) == 0);
            accept = (combined == 0) && trivial;
            nontrivial = (combined == 0) && !trivial && V.clipping;
        }
        const uint32_t packet = t >> 4, plane = t & 15;
        TriRecord rec;
        float wt[4];
        bool keep = false;
        if (accept) {
            keep = setup_triangle(AT, vv, hw, hh, cull_mode, rec, wt);
            rec.rank = make_rank(seq, packet, 0, plane, 0);
            rec.flags = (alpha ? kRecAlpha : 0u) | flags_hi;
        }
        unsigned long long km = __ballot(keep);
        if (keep) {
            const uint32_t slot = kept + mbcnt(km);
            out[slot] = rec;
            if (alpha) {
                AlphaRecord a;
                a.w0 = wt[0]; a.w0s = wt[1]; a.w1s = wt[2]; a.w2s = wt[3];
#pragma unroll
                for (int k = 0; k < 3; k++)
                    a.uv[k] = V.m.texcoords[(size_t)f.mi * 64 + ((f.idx[round] >> (8 * k)) & 63u)];
                a.tex = alpha_tex;
#pragma unroll
                for (int k = 0; k < 3; k++) a.ru[k] = a.rv[k] = 0.0f;
                a.pad[0] = a.pad[1] = 0;
                p.alpha[(size_t)i * 128 + slot] = a;
            }
        }
        kept += (uint32_t)__popcll(km);
        unsigned long long cm = __ballot(nontrivial);
        if (cm) {
            uint32_t n = __popcll(cm);
            clipped += n;
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&p.scratch->num_clip, n);
            base = __shfl(base, 0);
            const uint32_t info = partial | ((uint32_t)cull_mode << 8) | ((alpha ? 1u : 0u) << 10);
            const bool queued = nontrivial && base + mbcnt(cm) < p.clip_capacity;
            if (queued) {
                ClipEntry ce;
                ce.rank_base = make_rank(seq, packet, 1, plane, 0);
                ce.info = info;
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int c = 0; c < 4; c++) ce.v[k][c] = vv[k][c];
                ce.view = v;
                ce.pad = 0;
                p.clip_queue[base + mbcnt(cm)] = ce;
            }
            // Triangles that found the queue full go into free slots at the
            // top of the meshlet's own record block, to be clipped after the
            // binning (pipelines without that pass count them as dropped).
            const unsigned long long om = __ballot(nontrivial && !queued);
            if (om) {
                if (!p.exact_overflow) {
                    if (lane == 0) atomicAdd(&p.counters->dropped, (unsigned long long)__popcll(om));
                } else {
                    if (nontrivial && !queued) {
                        ClipRef r;
                        r.rank_base = make_rank(seq, packet, 1, plane, 0);
                        r.info = info;
                        r.meshlet = f.mi;
                        r.tri = t;
                        r.view = v;
                        r.pad[0] = r.pad[1] = r.pad[2] = 0;
                        reinterpret_cast<ClipRef*>(out)[127 - (overflowed + mbcnt(om))] = r;
                    }
                    overflowed += (uint32_t)__popcll(om);
                }
            }
        }
    }
    if (lane == 0) {
        p.meshlet_counts[i] = kept;
        if (overflowed) {
            // listed once; at most one entry per visible meshlet
            const uint32_t k = atomicAdd(&p.scratch->num_clip_overflow, 1u);
            p.clip_overflow[k] = make_uint2(i, overflowed);
        }
        if (counting(p)) {
            atomicAdd(&p.counters->triangles_processed, (unsigned long long)nt);
            atomicAdd(&p.counters->triangles_rasterized, (unsigned long long)kept);
            if (clipped) atomicAdd(&p.counters->triangles_clipped, (unsigned long long)clipped);
        }
    }
}

template <bool Alpha>
__global__ void __launch_bounds__(256) k_mesh_setup(DrawParams p) {
    __shared__ float4 lds_pos[4][64];
    // the VRCP14 / VRSQRT14 runs in LDS: triangle setup's reciprocals wait
    // on an LDS read instead of a cached global one
    __shared__ uint2 s_pwl[128];
    if (threadIdx.x < 128) s_pwl[threadIdx.x] = p.tables.pwl[threadIdx.x];
    __syncthreads();
    const ApproxTables lds_tables{ s_pwl };
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t nvis = p.scratch->num_visible;
    const uint32_t waves = gridDim.x * 4;
    uint32_t i = blockIdx.x * 4 + wid;
    if (i >= nvis) return;
    MeshletFetch cur = fetch_meshlet(p, i, lane);
    for (;;) {
        const uint32_t next = i + waves;
        MeshletFetch nxt;
        if (next < nvis) nxt = fetch_meshlet(p, next, lane);
        setup_meshlet<Alpha>(p, lds_tables, i, lane, lds_pos[wid], cur);
        if (next >= nvis) break;
        cur = nxt;
        i = next;
    }
}

// ---------------------------------------------------------------- clipper

// Clipper::ClipTriangles for one triangle (Rasterizer.cpp:398-478), one wave
// per queued triangle with the polygon spread over the lanes (lane i holds
// vertex i, in polygon order). For each plane set in the triangle's combined
// outcode, in plane order L, R, T, B, N, F: every lane computes its vertex's
// distance `+-c + w*scale` (guard-band scale for x/y), sees the next
// vertex's through a lane shuffle, emits itself when inside and the edge
// intersection `lerp(a, b, da/(da-db))` when the edge crosses; an exclusive
// scan of the emit counts gives each output's position and every output lane
// pulls its value from its source lane. Values and order are those of the
// reference's sequential loop. Generated vertices are limited to 16 per
// triangle, the room the reference's vertex store has. The fan triangles
// (v0, v[f+1], v[f+2]) then go through setup with culling (FlushPacket,
// Rasterizer.cpp:479-491) and are appended after the meshlet record slots.
__device__ __forceinline__ float lane_get(float v, uint32_t src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute((int)(src << 2), __float_as_int(v)));
}

template <bool Alpha>
__device__ void raster_now(const DrawParams& p, const TriRecord& rec, const AlphaRecord& A, uint32_t lane);

// Clips one triangle with the whole wave. With `Store`, lanes whose fan
// triangle survives setup return true with the record and its index, the
// record stored after the meshlet slots (the region holds 8 records per
// queue entry, as many as one triangle can yield, so it cannot fill up):
// packed through a shared counter, or, given `slots`, at slots + fan index
// (the entry's own 8 slots, no counter).
// Without `Store` (triangles that found the queue full) the surviving fan
// triangles are drawn at once by the wave (raster_now).
// With `Alpha` the barycentric attributes travel with the positions.
template <bool Alpha, bool Store>
__device__ bool clip_one(const DrawParams& p, const ClipSrc& ce, uint32_t lane, TriRecord& out,
                         uint32_t& out_index, uint32_t slots = 0xFFFFFFFFu) {
    constexpr int NC = Alpha ? 6 : 4;   // clipped channels
    const ViewParams& VP = view_of(p, ce.view);
    const int hw = VP.width / 2, hh = VP.height / 2;
    // x, y, z, w and the barycentric attributes (u, v) of the source
    // triangle's vertices, (0,0), (1,0), (0,1) (Rasterizer.cpp:399-408)
    float V[NC];
#pragma unroll
    for (int c = 0; c < 4; c++) V[c] = lane == 0 ? ce.v[0][c] : (lane == 1 ? ce.v[1][c] : (lane == 2 ? ce.v[2][c] : 0.0f));
    if constexpr (Alpha) {
        V[4] = lane == 1 ? 1.0f : 0.0f;
        V[5] = lane == 2 ? 1.0f : 0.0f;
    }
    uint32_t vc = 3, generated = 0;
    const uint32_t codes = ce.info & 0x3Fu;
    for (uint32_t plane = 0; plane < 6; plane++) {
        if (!((codes >> plane) & 1u)) continue;
        const float scale = plane < 2 ? VP.guard_x : (plane < 4 ? VP.guard_y : 1.0f);
        const uint32_t comp = plane >> 1;
        const bool neg = plane & 1u;
        const bool active = lane < vc;
        const uint32_t nxt = (lane + 1 < vc) ? lane + 1 : 0;
        float Vn[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) Vn[c] = lane_get(V[c], nxt);
        float ca = comp == 0 ? V[0] : (comp == 1 ? V[1] : V[2]);
        float cb = comp == 0 ? Vn[0] : (comp == 1 ? Vn[1] : Vn[2]);
        if (neg) { ca = -ca; cb = -cb; }
        const float da = __builtin_fmaf(V[3], scale, ca);
        const float db = __builtin_fmaf(Vn[3], scale, cb);
        const bool in = active && da >= 0.0f;
        const bool cross = active && ((da >= 0.0f) != (db >= 0.0f));
        // crossings are created in polygon order while room is left
        const unsigned long long xm = __ballot(cross);
        const uint32_t xrank = mbcnt(xm);
        const bool make = cross && (generated + xrank < 16);
        const uint32_t made = min((uint32_t)__popcll(xm), 16u - generated);
        const uint32_t cnt = (in ? 1u : 0u) + (make ? 1u : 0u);
        const uint32_t incl = wave_scan_add(cnt);
        const uint32_t outCount = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t pos = incl - cnt;
        float I[NC] = {};
        if (make) {
            const float t = da / (da - db);
#pragma unroll
            for (int c = 0; c < NC; c++) I[c] = __builtin_fmaf(Vn[c], t, __builtin_fmaf(-t, V[c], V[c]));
        }
        if (vc < 3) break;
        // output lane j pulls from source lane src(j) = #{i : incl_i <= j}
        uint32_t src = 0;
        for (uint32_t k = 0; k < vc; k++) src += ((uint32_t)__builtin_amdgcn_readlane((int)incl, k) <= lane) ? 1u : 0u;
        src = min(src, 63u);
        const uint32_t srcPos = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)pos);
        const int srcIn = __builtin_amdgcn_ds_bpermute((int)(src << 2), in ? 1 : 0);
        const bool isVert = srcIn && lane == srcPos;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const float a = lane_get(V[c], src), b = lane_get(I[c], src);
            V[c] = isVert ? a : b;
        }
        generated += made;
        vc = outCount;
    }
    if (vc < 3) return false;
    const int cull_mode = (ce.info >> 8) & 3;
    // fan triangle f = (v0, v[f+1], v[f+2]) on lane f
    const uint32_t f = lane;
    float tri[3][4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        tri[0][c] = lane_get(V[c], 0);
        tri[1][c] = lane_get(V[c], min(f + 1, 63u));
        tri[2][c] = lane_get(V[c], min(f + 2, 63u));
    }
    // TriangleClipData of the fan triangle (Rasterizer.cpp:461-464)
    float u0 = 0, v0 = 0, ru1 = 0, rv1 = 0, ru2 = 0, rv2 = 0;
    if constexpr (Alpha) {
        u0 = lane_get(V[4], 0);
        v0 = lane_get(V[5], 0);
        ru1 = lane_get(V[4], min(f + 1, 63u)) - u0;
        rv1 = lane_get(V[5], min(f + 1, 63u)) - v0;
        ru2 = lane_get(V[4], min(f + 2, 63u)) - u0;
        rv2 = lane_get(V[5], min(f + 2, 63u)) - v0;
    }
    const bool alpha = Alpha && ((ce.info >> 10) & 1u);
    TriRecord rec;
    float wt[4];
    bool keep = false;
    if (f + 2 < vc && f < 8) {
        keep = setup_triangle(p.tables, tri, hw, hh, cull_mode, rec, wt);
        rec.rank = ce.rank_base | f;
        rec.flags = (alpha ? kRecAlpha : 0u) | (ce.view << 16);
    }
    const unsigned long long km = __ballot(keep);
    if (!km) return false;
    if (lane == 0 && counting(p)) atomicAdd(&p.counters->triangles_rasterized, (unsigned long long)__popcll(km));
    AlphaRecord a = {};
    if (alpha && keep) {
        // the source triangle's UVs and material, from the meshlet and
        // triangle the rank names
        const uint32_t seq = ce.rank_base >> 11;
        const uint32_t t = ((ce.rank_base >> 8) & 7u) * 16 + ((ce.rank_base >> 3) & 15u);
        const uint32_t mi = VP.meshlet_offset + (seq - VP.seq_base);
        const DevMaterial dm = VP.materials[VP.m.material_id[mi]];
        a.w0 = wt[0]; a.w0s = wt[1]; a.w1s = wt[2]; a.w2s = wt[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            a.uv[k] = VP.m.texcoords[(size_t)mi * 64 + (VP.m.indices[(size_t)mi * 384 + k * 128 + t] & 63u)];
        a.tex = (uint32_t)dm.texture | ((uint32_t)dm.alpha_cutoff << 24);
        a.ru[0] = u0; a.ru[1] = ru1; a.ru[2] = ru2;
        a.rv[0] = v0; a.rv[1] = rv1; a.rv[2] = rv2;
    }
    if constexpr (Store) {
        uint32_t slot = slots + f;
        if (slots == 0xFFFFFFFFu) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&p.scratch->num_records, (uint32_t)__popcll(km));
            slot = (uint32_t)__builtin_amdgcn_readlane((int)base, 0) + mbcnt(km);
        }
        if (!keep) return false;
        if (slot >= p.clip_record_capacity) {     // not reached: see above
            atomicAdd(&p.counters->dropped, 1ull);
            return false;
        }
        p.records[p.clip_record_base + slot] = rec;
        if (alpha) p.alpha[p.clip_record_base + slot] = a;
        out = rec;
        out_index = p.clip_record_base + slot;
        return true;
    } else {
        // one fan triangle at a time, broadcast to the wave
        unsigned long long nm = km;
        while (nm) {
            const int l = __ffsll((long long)nm) - 1;
            nm &= nm - 1;
            TriRecord r;
#pragma unroll
            for (int k = 0; k < 3; k++) r.pos[k] = (uint32_t)__shfl((int)rec.pos[k], l);
            r.z0 = __shfl(rec.z0, l);
            r.z10 = __shfl(rec.z10, l);
            r.z20 = __shfl(rec.z20, l);
            r.rank = (uint32_t)__shfl((int)rec.rank, l);
            r.flags = (uint32_t)__shfl((int)rec.flags, l);
            AlphaRecord ar = {};
            if constexpr (Alpha) {
                if (r.flags & kRecAlpha) {
                    ar.w0 = __shfl(a.w0, l); ar.w0s = __shfl(a.w0s, l);
                    ar.w1s = __shfl(a.w1s, l); ar.w2s = __shfl(a.w2s, l);
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        ar.uv[k] = (uint32_t)__shfl((int)a.uv[k], l);
                        ar.ru[k] = __shfl(a.ru[k], l);
                        ar.rv[k] = __shfl(a.rv[k], l);
                    }
                    ar.tex = (uint32_t)__shfl((int)a.tex, l);
                }
            }
            raster_now<Alpha>(p, r, ar, lane);
        }
        return false;
    }
}

// The triangles setup could not queue (ClipRef): one wave per listed
// meshlet clips its references in order and draws their fan triangles at
// once. Runs after setup, before the bin raster.
template <bool Alpha>
__device__ void clip_overflow(const DrawParams& p, uint32_t first_wave, uint32_t waves) {
    const uint32_t lane = lane_id();
    const uint32_t n = p.scratch->num_clip_overflow;
    for (uint32_t k = first_wave; k < n; k += waves) {
        const uint2 e = p.clip_overflow[k];      // (visible slot, references)
        const ClipRef* refs = reinterpret_cast<const ClipRef*>(p.records + (size_t)e.x * 128);
        for (uint32_t j = 0; j < e.y; j++) {
            const ClipRef r = refs[127 - j];
            const ViewParams& V = view_of(p, r.view);
            ClipSrc cs;
            cs.rank_base = r.rank_base;
            cs.info = r.info;
            cs.view = r.view;
            // the vertices as setup transformed them (ShadeMeshlet and the
            // gather_preload64 wrap of the index)
            const float* P = V.m.positions + (size_t)r.meshlet * 192;
            const uint8_t* I = V.m.indices + (size_t)r.meshlet * 384;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const uint32_t ix = I[k * 128 + r.tri];
                const uint32_t vi = ix < 32 ? ix : 32 + (ix & 31);
#pragma unroll
                for (int c = 0; c < 4; c++) cs.v[k][c] = xform(V.clip, c, P[vi], P[64 + vi], P[128 + vi]);
            }
            TriRecord unused;
            uint32_t unused_index;
            (void)clip_one<Alpha, false>(p, cs, lane, unused, unused_index);
        }
    }
}

template <bool Alpha>
__global__ void __launch_bounds__(256) k_clip(DrawParams p) {
    const uint32_t n = min(p.scratch->num_clip, p.clip_capacity);
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t q = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; q < n; q += waves) {
        TriRecord rec;
        uint32_t idx;
        (void)clip_one<Alpha, true>(p, clip_src(p.clip_queue[q]), lane, rec, idx);
    }
    // then the triangles that found the queue full (direct route)
    clip_overflow<Alpha>(p, blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, waves);
}


// ---------------------------------------------------------------- record chunks

// Chunks of up to 64 records for a wave: chunk c < 2*num_visible is half
// c&1 of visible meshlet slot c>>1; later chunks cover the clipped records.
struct Chunk {
    uint32_t base, count;
    const uint32_t* list;      // non-null: record index = list[base + i]
};

__device__ __forceinline__ uint32_t num_chunks(const DrawParams& p, bool with_clipped = true) {
    if (!with_clipped) return 2 * p.scratch->num_visible;
    uint32_t nclip = min(p.scratch->num_records, p.clip_record_capacity);
    return 2 * p.scratch->num_visible + (nclip + 63) / 64;
}

__device__ __forceinline__ Chunk get_chunk(const DrawParams& p, uint32_t c) {
    Chunk k;
    k.list = nullptr;
    const uint32_t nvis = p.scratch->num_visible;
    if (c < 2 * nvis) {
        uint32_t i = c >> 1, half = c & 1;
        uint32_t cnt = p.meshlet_counts[i];
        k.base = i * 128 + half * 64;
        k.count = cnt > half * 64 ? min(64u, cnt - half * 64) : 0;
        return k;
    }
    uint32_t nclip = min(p.scratch->num_records, p.clip_record_capacity);
    uint32_t off = (c - 2 * nvis) * 64;
    k.base = p.clip_record_base + off;
    k.count = min(64u, nclip - off);
    return k;
}

// ---------------------------------------------------------------- direct raster

// Per-triangle raster state kept in LDS while a wave or block walks the
// triangles' 4x4 tiles.
struct TriLds {
    int32_t e0, e1, e2, a12, a20, a01, b12, b20, b01;
    float z0, z10, z20;
    uint32_t keylo, org, tw, start;
};

// The bin raster's state of one record, prepared once per (record, bin) so
// that a 16-lane group stepping onto the record pays only a few multiply-adds:
// the edge values at the first pixel of the record's render box clipped to
// the bin, the edge steps, the depth plane, the key's low word, the LDS key
// index of the box's first tile and the record's first tile in the block's
// tile list. In bins of up to 32x32 pixels (at most 8x8 tiles) only the
// box's tiles that the triangle can cover are walked, the set bits of a
// per-record tile mask; 64-pixel bins walk every tile of the box, with the
// step from the last tile of a tile row to the first of the next
// (4B - 4A (tw - 1)).
struct BinTriMask {
    int32_t w0, w1, w2;
    int32_t a12, a20, a01, b12, b20, b01;
    float z0, z10, z20;
    uint32_t keylo, li0, start, pad;
};
static_assert(sizeof(BinTriMask) == 64, "BinTriMask is four 16-byte LDS reads");

struct BinTriStep {
    int32_t w0, w1, w2;
    int32_t a12, a20, a01, b12, b20, b01;
    int32_t r0, r1, r2;
    float z0, z10, z20;
    uint32_t keylo, li0, tw, start, pad;
};
static_assert(sizeof(BinTriStep) == 80, "BinTriStep is five 16-byte LDS reads");

template <int BS>
using BinTri = typename std::conditional<(BS <= 32), BinTriMask, BinTriStep>::type;

// Entries of the bin raster's per-record tile masks in LDS (one per thread
// for bins of up to 32 pixels; the 64-pixel walk has none)
__host__ __device__ constexpr uint32_t bin_masks(int BS, int kThreads) {
    return BS <= 32 ? (uint32_t)kThreads : 1u;
}

__device__ __forceinline__ void load_tri(const TriRecord& rec, int hw, int hh, TriLds& t, uint32_t& bmin,
                                         uint32_t& bmax) {
    render_bbox(rec.pos[0], rec.pos[1], rec.pos[2], hw, hh, bmin, bmax);
    Edges e = edge_setup(rec.pos[0], rec.pos[1], rec.pos[2], hw, hh);
    t.e0 = e.e0; t.e1 = e.e1; t.e2 = e.e2;
    t.a12 = e.a12; t.a20 = e.a20; t.a01 = e.a01;
    t.b12 = e.b12; t.b20 = e.b20; t.b01 = e.b01;
    t.z0 = rec.z0; t.z10 = rec.z10; t.z20 = rec.z20;
    t.keylo = ~rec.rank;
}

// Coverage and depth of pixel (x, y) for triangle T (DrawTriangle,
// Rasterizer.h:281-296). Returns false when the pixel is not covered.
// Edge steps are differences of s16 coordinates (|A|, |B| < 2^17) and pixel
// coordinates are below 2^12, so full-rate 24-bit multiplies give the exact
// products; the sums wrap in 32 bits like the reference's.
__device__ __forceinline__ int32_t edge_at(int32_t e, int32_t a, int32_t b, uint32_t x, uint32_t y) {
    return (int32_t)((uint32_t)e + (uint32_t)__mul24(a, (int32_t)x) + (uint32_t)__mul24(b, (int32_t)y));
}

__device__ __forceinline__ bool shade_pixel(const TriLds& T, uint32_t x, uint32_t y, float& depth) {
    int32_t w0 = edge_at(T.e0, T.a12, T.b12, x, y);
    int32_t w1 = edge_at(T.e1, T.a20, T.b20, x, y);
    int32_t w2 = edge_at(T.e2, T.a01, T.b01, x, y);
    if ((w0 | w1 | w2) < 0) return false;
    depth = __builtin_fmaf((float)w1, T.z10, __builtin_fmaf((float)w2, T.z20, T.z0));
    return true;
}

// UVs of one fragment of a 4x4 tile as FS_EncodeSurfaceId<true> and
// FS_EncodeGBuffer compute them (lane & 15 = x + 4y inside the tile). All 16
// lanes must be active: the screen derivatives and the filter choice are
// taken across them. (e1, e2) are the lane's edge values, the u and v of
// DrawTriangle: perspective-correct barycentrics (Rasterizer.h:298-314),
// the clipped triangle's remap (Rasterizer.h:306-312), BaryLerp of the
// vertex UVs and SampleImplicitLod's level (Texture.h:403-410: derivatives
// from the tile's pixel pairs, mip level less the 8 fraction bits,
// minification anywhere in the tile selects the min filter). Operation
// order follows the reference's optimised build: `u *= W1S * rcpW` is
// evaluated as (u * W1S) * rcpW and `1 - u - v` as 1 - (u + v).
struct TileUV {
    float U, V;
    int32_t mip;
    bool minify;
    float b[3];                    // Bary = {1 - u - v, u, v}
};

__device__ __forceinline__ TileUV tile_uv(const ApproxTables& T, const AlphaRecord& A, bool clipped,
                                          const DevTexture& tex, int32_t e1, int32_t e2, uint32_t lane) {
    const float u = (float)e1, v = (float)e2;
    const float pw0 = __builtin_fmaf(u + v, -A.w0s, A.w0);
    const float w = __builtin_fmaf(u, A.w1s, __builtin_fmaf(v, A.w2s, pw0));
    float r = rcp14(T, w);
    r = r * __builtin_fmaf(-w, r, 2.0f);
    float bu = (u * A.w1s) * r, bv = (v * A.w2s) * r;
    if (clipped) {
        const float cu = __builtin_fmaf(bu, A.ru[1], __builtin_fmaf(bv, A.ru[2], A.ru[0]));
        const float cv = __builtin_fmaf(bu, A.rv[1], __builtin_fmaf(bv, A.rv[2], A.rv[0]));
        bu = cu;
        bv = cv;
    }
    TileUV o;
    const float b0 = 1.0f - (bu + bv);
    o.b[0] = b0;
    o.b[1] = bu;
    o.b[2] = bv;
    float tu[3], tv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        tu[k] = half_to_float(A.uv[k] & 0xFFFFu);
        tv[k] = half_to_float(A.uv[k] >> 16);
    }
    o.U = __builtin_fmaf(tu[0], b0, __builtin_fmaf(tu[1], bu, tu[2] * bv));
    o.V = __builtin_fmaf(tv[0], b0, __builtin_fmaf(tv[1], bu, tv[2] * bv));
    const float su = o.U * tex.scale_lerp_u, sv = o.V * tex.scale_lerp_v;
    // dFdx pairs pixels (2i, 2i+1) of a row, dFdy rows (r, r+1) of a pair
    const float sux = __shfl_xor(su, 1), svx = __shfl_xor(sv, 1);
    const float suy = __shfl_xor(su, 4), svy = __shfl_xor(sv, 4);
    const bool odd = (lane & 1u) != 0, low = (lane & 4u) != 0;
    const float gx = odd ? su - sux : sux - su, gy = odd ? sv - svx : svx - sv;
    const float gz = low ? su - suy : suy - su, gw = low ? sv - svy : svy - sv;
    const float dx = __builtin_fmaf(gx, gx, gy * gy), dy = __builtin_fmaf(gz, gz, gw * gw);
    const float mx = (dx > dy) ? dx : dy;
    o.mip = ((int32_t)(fbits(mx) - (127u << 23)) >> 24) - 8;
    const unsigned long long groupMask = 0xFFFFull << (lane & 48u);
    o.minify = (__ballot(o.mip > 0) & groupMask) != 0;
    return o;
}

// The alpha test of FS_EncodeSurfaceId<true> (Shading.cpp:316-327) for the
// 16 lanes of one 4x4 tile: SampleImplicitLod at the fragment's UV, then
// color >= cutoff << 24.
__device__ __forceinline__ bool alpha_test_tile(const ApproxTables& T, const AlphaRecord& A, bool clipped,
                                                const DevTexture* textures, int32_t e1, int32_t e2, uint32_t lane) {
    const DevTexture& tex = textures[A.tex & 0xFFFFFFu];
    const TileUV t = tile_uv(T, A, clipped, tex, e1, e2, lane);
    const uint32_t color = sample_level(tex, t.U, t.V, t.mip, !t.minify);
    return color >= ((A.tex >> 24) << 24);
}

// Draws one record with the whole wave, four 4x4 tiles at a time (lane =
// pixel), for a clipped fan triangle that found no record slot: into the
// view's spill keys when the launch bins (marking each bin the record passes
// a fragment in, so that bin's raster merges its spill keys) or straight
// into the vis-buffer on the direct route. `rec` and `A` are wave-uniform.
template <bool Alpha>
__device__ void raster_now(const DrawParams& p, const TriRecord& rec, const AlphaRecord& A, uint32_t lane) {
    const ViewParams& V = view_of(p, rec.flags >> 16);
    const uint32_t W = (uint32_t)V.width;
    TriLds T;
    uint32_t bmin, bmax;
    load_tri(rec, V.width / 2, V.height / 2, T, bmin, bmax);
    const uint32_t x0 = (uint32_t)lo16(bmin), y0 = (uint32_t)hi16(bmin);
    const uint32_t tw = ((uint32_t)lo16(bmax) - x0) >> 2, total = tw * (((uint32_t)hi16(bmax) - y0) >> 2);
    const bool binned = p.total_bins != 0;
    unsigned long long* dst_keys = binned ? V.spill_keys : V.keys;
    const bool alpha = Alpha && (rec.flags & kRecAlpha) != 0;
    const float clearDepth = V.clear_depth;        // in a register: see raster_bin
    const uint32_t sub = lane >> 4, px = lane & 3, py = (lane >> 2) & 3;
    uint32_t covered = 0, passed = 0;
    for (uint32_t base = 0; base < total; base += 4) {
        const uint32_t j = base + sub;
        if (j >= total) continue;                    // uniform across each 16-lane group
        const uint32_t ty = j / tw, tx = j - ty * tw;
        const uint32_t x = x0 + tx * 4 + px, y = y0 + ty * 4 + py;
        const int32_t w0 = edge_at(T.e0, T.a12, T.b12, x, y);
        const int32_t w1 = edge_at(T.e1, T.a20, T.b20, x, y);
        const int32_t w2 = edge_at(T.e2, T.a01, T.b01, x, y);
        const bool cov = (w0 | w1 | w2) >= 0;
        const float depth = __builtin_fmaf((float)w1, T.z10, __builtin_fmaf((float)w2, T.z20, T.z0));
        bool pass = cov && depth > clearDepth;
        if constexpr (Alpha) {
            if (alpha && (__ballot(pass) & (0xFFFFull << (lane & 48u)))) {
                const bool ok = alpha_test_tile(p.tables, A, ((T.keylo >> 7) & 1u) == 0, V.textures, w1, w2, lane);
                pass = pass && ok;
            }
        }
        if (cov) covered++;
        if (pass) {
            passed++;
            const unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
            unsigned long long* dst = dst_keys + pixel_offset(W, x, y);
            if (key > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, key);
        }
        // the ballot is taken by every lane of the group (a ballot behind the
        // lane test would see the group's first lane only)
        const unsigned long long passing = __ballot(pass) & (0xFFFFull << (lane & 48u));
        if (binned && (lane & 15u) == 0 && passing)
            p.bin_flags[V.bin_base + ((y0 + ty * 4) >> p.bin_shift) * V.bins_x + ((x0 + tx * 4) >> p.bin_shift)] = 1u;
    }
    if (counting(p)) {
        for (int d = 32; d > 0; d >>= 1) {
            covered += __shfl_xor(covered, d);
            passed += __shfl_xor(passed, d);
        }
        if (lane == 0) {
            if (covered) atomicAdd(&p.counters->fragments_covered, (unsigned long long)covered);
            if (passed) atomicAdd(&p.counters->fragments_passed, (unsigned long long)passed);
        }
    }
}

// Each wave takes a chunk of up to 64 records (meshlet slots or clipped
// records), lays their render boxes out as one list of
// 4x4 tiles (exclusive scan of tile counts) and walks it four tiles at a
// time: lane = pixel, 16 lanes = one tile, so a tile's 16 keys (128 bytes)
// are one contiguous access. Covered pixels update the key of their view's
// vis-buffer with atomicMax when the key can still win. With `Alpha`,
// tiles of alpha-tested records that have a fragment passing the depth test
// run the alpha test (the reference tests after the depth test; the test is
// a function of the triangle and the tile alone, so the order does not
// change which fragments pass).
template <bool Alpha>
__global__ void __launch_bounds__(256) k_raster_direct(DrawParams p) {
    __shared__ TriLds tl[4][64];
    __shared__ uint32_t tview[4][64];
    __shared__ uint32_t talpha[Alpha ? 4 : 1][64];   // record index of an alpha-tested record, else ~0
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t nchunks = num_chunks(p);
    const bool countFrags = counting(p);
    const uint32_t waves = gridDim.x * 4;

    for (uint32_t c = blockIdx.x * 4 + wid; c < nchunks; c += waves) {
        const Chunk ck = get_chunk(p, c);
        if (ck.count == 0) continue;
        uint32_t ntiles = 0;
        TriLds t;
        if (lane < ck.count) {
            const uint32_t r = ck.list ? ck.list[ck.base + lane] : ck.base + lane;
            const TriRecord rec = p.records[r];
            const uint32_t v = rec.flags >> 16;
            const ViewParams& V = view_of(p, v);
            uint32_t bmin, bmax;
            load_tri(rec, V.width / 2, V.height / 2, t, bmin, bmax);
            t.org = bmin;
            uint32_t tw = (uint32_t)(lo16(bmax) - lo16(bmin)) >> 2;
            uint32_t th = (uint32_t)(hi16(bmax) - hi16(bmin)) >> 2;
            t.tw = tw;
            ntiles = tw * th;
            tview[wid][lane] = v;
            if constexpr (Alpha) talpha[wid][lane] = (rec.flags & kRecAlpha) ? r : 0xFFFFFFFFu;
        }
        const uint32_t incl = wave_scan_add(ntiles);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        t.start = incl - ntiles;
        if (lane < ck.count) tl[wid][lane] = t;
        else tl[wid][lane].start = 0xFFFFFFFFu;
        __builtin_amdgcn_wave_barrier();

        uint32_t covered = 0, passed = 0;
        const uint32_t sub = lane >> 4, px = lane & 3, py = (lane >> 2) & 3;
        for (uint32_t base = 0; base < total; base += 4) {
            const uint32_t j = base + sub;
            if (j < total) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1) {
                    if (tl[wid][k + step].start <= j) k += step;
                }
                const TriLds& T = tl[wid][k];
                const ViewParams& V = view_of(p, tview[wid][k]);
                const uint32_t W = (uint32_t)V.width;
                const uint32_t pt = j - T.start;
                const uint32_t ty = pt / T.tw;
                const uint32_t tx = pt - ty * T.tw;
                const uint32_t x = (uint32_t)lo16(T.org) + tx * 4 + px;
                const uint32_t y = (uint32_t)hi16(T.org) + ty * 4 + py;
                if constexpr (!Alpha) {
                    float depth;
                    if (x < W && y < (uint32_t)V.height && shade_pixel(T, x, y, depth)) {
                        covered++;
                        if (depth > V.clear_depth) {
                            unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
                            unsigned long long* dst = V.keys + pixel_offset(W, x, y);
                            if (key > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, key);
                            passed++;
                        }
                    }
                } else {
                    // every lane of the tile evaluates its edges (the alpha
                    // test needs all 16); render boxes lie inside the
                    // framebuffer, the bounds test only guards the store
                    const int32_t w0 = edge_at(T.e0, T.a12, T.b12, x, y);
                    const int32_t w1 = edge_at(T.e1, T.a20, T.b20, x, y);
                    const int32_t w2 = edge_at(T.e2, T.a01, T.b01, x, y);
                    const bool cov = x < W && y < (uint32_t)V.height && (w0 | w1 | w2) >= 0;
                    const float depth = __builtin_fmaf((float)w1, T.z10, __builtin_fmaf((float)w2, T.z20, T.z0));
                    bool pass = cov && depth > V.clear_depth;
                    const uint32_t ai = talpha[wid][k];
                    if (ai != 0xFFFFFFFFu && (__ballot(pass) & (0xFFFFull << (lane & 48u)))) {
                        const bool ok = alpha_test_tile(p.tables, p.alpha[ai], ((T.keylo >> 7) & 1u) == 0,
                                                        V.textures, w1, w2, lane);
                        pass = pass && ok;
                    }
                    if (cov) covered++;
                    if (pass) {
                        unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
                        unsigned long long* dst = V.keys + pixel_offset(W, x, y);
                        if (key > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, key);
                        passed++;
                    }
                }
            }
        }
        if (countFrags) {
            for (int d = 32; d > 0; d >>= 1) {
                covered += __shfl_xor(covered, d);
                passed += __shfl_xor(passed, d);
            }
            if (lane == 0) {
                atomicAdd(&p.counters->fragments_covered, (unsigned long long)covered);
                atomicAdd(&p.counters->fragments_passed, (unsigned long long)passed);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The overdraw pipeline: Rasterizer::DrawTriangle with FS_Overdraw
// (Rasterizer.h:281-324, Shading.cpp:333-342). Every 4x4 tile of a
// triangle's render box with a covered pixel adds one to each of its 16
// pixels: to the covered count when the pixel is covered, else to the
// quad-helper count; and raises each pixel's depth to the triangle's plane
// depth there (extrapolated for uncovered pixels) when that is greater.
// Counts and maxima commute, so the result does not depend on the order in
// which triangles are drawn. Counts are 32-bit halves of one 64-bit word
// per pixel (saturated to 16 bits when read); the depth maximum is a 32-bit
// max on the float's bits, taken only for positive depths, which order like
// their bits (the stored depth starts at the clear depth, >= +0).
__global__ void __launch_bounds__(256) k_raster_overdraw(DrawParams p) {
    __shared__ TriLds tl[4][64];
    __shared__ uint32_t tview[4][64];
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t nchunks = num_chunks(p);
    const bool countFrags = counting(p);
    const uint32_t waves = gridDim.x * 4;
    const unsigned long long groupMask = 0xFFFFull << (lane & 48u);

    for (uint32_t c = blockIdx.x * 4 + wid; c < nchunks; c += waves) {
        const Chunk ck = get_chunk(p, c);
        if (ck.count == 0) continue;
        uint32_t ntiles = 0;
        TriLds t;
        if (lane < ck.count) {
            const uint32_t r = ck.list ? ck.list[ck.base + lane] : ck.base + lane;
            const TriRecord rec = p.records[r];
            const uint32_t v = rec.flags >> 16;
            const ViewParams& V = view_of(p, v);
            uint32_t bmin, bmax;
            load_tri(rec, V.width / 2, V.height / 2, t, bmin, bmax);
            t.org = bmin;
            const uint32_t tw = (uint32_t)(lo16(bmax) - lo16(bmin)) >> 2;
            const uint32_t th = (uint32_t)(hi16(bmax) - hi16(bmin)) >> 2;
            t.tw = tw;
            ntiles = tw * th;
            tview[wid][lane] = v;
        }
        const uint32_t incl = wave_scan_add(ntiles);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        t.start = incl - ntiles;
        if (lane < ck.count) tl[wid][lane] = t;
        else tl[wid][lane].start = 0xFFFFFFFFu;
        __builtin_amdgcn_wave_barrier();

        uint32_t covered = 0;
        const uint32_t sub = lane >> 4, px = lane & 3, py = (lane >> 2) & 3;
        for (uint32_t base = 0; base < total; base += 4) {
            const uint32_t j = base + sub;
            bool cov = false, inside = false;
            float depth = 0.0f;
            uint32_t off = 0;
            unsigned long long* keys = nullptr;
            uint32_t* dbuf = nullptr;
            if (j < total) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1) {
                    if (tl[wid][k + step].start <= j) k += step;
                }
                const TriLds& T = tl[wid][k];
                const ViewParams& V = view_of(p, tview[wid][k]);
                const uint32_t W = (uint32_t)V.width;
                const uint32_t pt = j - T.start;
                const uint32_t ty = pt / T.tw;
                const uint32_t tx = pt - ty * T.tw;
                const uint32_t x = (uint32_t)lo16(T.org) + tx * 4 + px;
                const uint32_t y = (uint32_t)hi16(T.org) + ty * 4 + py;
                const int32_t w0 = edge_at(T.e0, T.a12, T.b12, x, y);
                const int32_t w1 = edge_at(T.e1, T.a20, T.b20, x, y);
                const int32_t w2 = edge_at(T.e2, T.a01, T.b01, x, y);
                // render boxes lie inside the framebuffer; the bounds test
                // only guards the stores
                inside = x < W && y < (uint32_t)V.height;
                cov = inside && (w0 | w1 | w2) >= 0;
                depth = __builtin_fmaf((float)w1, T.z10, __builtin_fmaf((float)w2, T.z20, T.z0));
                off = pixel_offset(W, x, y);
                keys = V.keys;
                dbuf = V.aux;
            }
            if (!(__ballot(cov) & groupMask) || !inside) continue;
            atomicAdd(keys + off, cov ? (1ull << 32) : 1ull);
            if (depth > 0.0f) {
                const uint32_t bits = fbits(depth);
                if (bits > __atomic_load_n(dbuf + off, __ATOMIC_RELAXED)) atomicMax(dbuf + off, bits);
            }
            covered += cov ? 1u : 0u;
        }
        if (countFrags) {
            for (int d = 32; d > 0; d >>= 1) covered += __shfl_xor(covered, d);
            if (lane == 0) {
                atomicAdd(&p.counters->fragments_covered, (unsigned long long)covered);
                atomicAdd(&p.counters->fragments_passed, (unsigned long long)covered);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// cvttps2udq: truncation; NaN and results outside [0, 2^32) give 0xFFFFFFFF.
__device__ __forceinline__ uint32_t cvt_trunc_u32(float x) {
    if (!(x > -1.0f && x < 4294967296.0f)) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// The layer-2 word of FS_EncodeGBuffer (Shading.cpp:397-401): MapOctahedron
// (Texture.h:282-288) of the remapped normal, each coordinate as
// conv<uint32_t>(oct * 1023 + 0.5) in 10 bits, then the NMR texel's
// metallic and roughness bytes cut to 6 bits. The optimised build folds
// (c * 0.5 + 0.5) * 1023 + 0.5 into fma(c, 511.5, 512), sums the absolute
// values left to right, takes 1/sum from VRCP14PS alone and evaluates
// max(-z * w, 0) as zero unless positive.
__device__ __forceinline__ uint32_t gbuffer_pack(const ApproxTables& T, const float n[3], uint32_t nmr) {
    const float w = rcp14(T, (fabsf(n[0]) + fabsf(n[1])) + fabsf(n[2]));
    const float mz = -n[2] * w;
    const float t = mz > 0.0f ? mz : 0.0f;
    const float u = __builtin_fmaf(n[0], w, mulsign(t, n[0]));
    const float v = __builtin_fmaf(n[1], w, mulsign(t, n[1]));
    uint32_t r = cvt_trunc_u32(__builtin_fmaf(u, 511.5f, 512.0f));
    r |= cvt_trunc_u32(__builtin_fmaf(v, 511.5f, 512.0f)) << 10;
    r |= ((nmr >> 18) & 0x3Fu) << 20;
    r |= ((nmr >> 26) & 0x3Fu) << 26;
    return r;
}

// FS_EncodeGBuffer's attribute stores (Shading.cpp:357-401) as a second pass
// over the draw's records, after the vis-buffer raster has settled which
// fragment wins each pixel (depth test, draw order and the alpha test,
// which FS_EncodeGBuffer applies to every material). A fragment wins when
// its key equals the pixel's final key; keys are unique per fragment, so
// exactly one fragment per covered pixel writes, and the result is the
// reference's last write. A tile with a winning fragment computes UVs and
// levels on all 16 lanes (derivatives and the filter choice span the tile),
// samples the base colour and, for two-layer textures, the NMR texel, and
// the winning lanes write the colour and the packed normal / metallic /
// roughness word (zero for one-layer textures). Meshlets of a deferred
// draw all have textured materials (checked on the host), so every record
// has its alpha record.
__global__ void __launch_bounds__(256) k_gbuffer(DrawParams p) {
    __shared__ TriLds tl[4][64];
    __shared__ uint32_t tview[4][64];
    __shared__ uint32_t talpha[4][64];
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t nchunks = num_chunks(p);
    const uint32_t waves = gridDim.x * 4;
    const unsigned long long groupMask = 0xFFFFull << (lane & 48u);
    const ApproxTables& AT = p.tables;

    for (uint32_t c = blockIdx.x * 4 + wid; c < nchunks; c += waves) {
        const Chunk ck = get_chunk(p, c);
        if (ck.count == 0) continue;
        uint32_t ntiles = 0;
        TriLds t;
        if (lane < ck.count) {
            const uint32_t r = ck.list ? ck.list[ck.base + lane] : ck.base + lane;
            const TriRecord rec = p.records[r];
            const uint32_t v = rec.flags >> 16;
            const ViewParams& V = view_of(p, v);
            uint32_t bmin, bmax;
            load_tri(rec, V.width / 2, V.height / 2, t, bmin, bmax);
            t.org = bmin;
            const uint32_t tw = (uint32_t)(lo16(bmax) - lo16(bmin)) >> 2;
            const uint32_t th = (uint32_t)(hi16(bmax) - hi16(bmin)) >> 2;
            t.tw = tw;
            ntiles = tw * th;
            tview[wid][lane] = v;
            talpha[wid][lane] = (rec.flags & kRecAlpha) ? r : 0xFFFFFFFFu;
        }
        const uint32_t incl = wave_scan_add(ntiles);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        t.start = incl - ntiles;
        if (lane < ck.count) tl[wid][lane] = t;
        else tl[wid][lane].start = 0xFFFFFFFFu;
        __builtin_amdgcn_wave_barrier();

        const uint32_t sub = lane >> 4, px = lane & 3, py = (lane >> 2) & 3;
        for (uint32_t base = 0; base < total; base += 4) {
            const uint32_t j = base + sub;
            bool match = false;
            uint32_t k = 0, off = 0;
            int32_t w1 = 0, w2 = 0;
            if (j < total) {
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1) {
                    if (tl[wid][k + step].start <= j) k += step;
                }
                const TriLds& T = tl[wid][k];
                const ViewParams& V = view_of(p, tview[wid][k]);
                const uint32_t W = (uint32_t)V.width;
                const uint32_t pt = j - T.start;
                const uint32_t ty = pt / T.tw;
                const uint32_t tx = pt - ty * T.tw;
                const uint32_t x = (uint32_t)lo16(T.org) + tx * 4 + px;
                const uint32_t y = (uint32_t)hi16(T.org) + ty * 4 + py;
                const int32_t w0 = edge_at(T.e0, T.a12, T.b12, x, y);
                w1 = edge_at(T.e1, T.a20, T.b20, x, y);
                w2 = edge_at(T.e2, T.a01, T.b01, x, y);
                const float depth = __builtin_fmaf((float)w1, T.z10, __builtin_fmaf((float)w2, T.z20, T.z0));
                if (x < W && y < (uint32_t)V.height && (w0 | w1 | w2) >= 0 && depth > V.clear_depth) {
                    off = pixel_offset(W, x, y);
                    const unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
                    match = V.keys[off] == key;
                }
            }
            // uniform per 16-lane group: the group shares j, hence the record
            if (!(__ballot(match) & groupMask)) continue;
            const uint32_t ai = talpha[wid][k];
            if (ai == 0xFFFFFFFFu) continue;
            const TriLds& T = tl[wid][k];
            const ViewParams& V = view_of(p, tview[wid][k]);
            const AlphaRecord& A = p.alpha[ai];
            const DevTexture& tex = V.textures[A.tex & 0xFFFFFFu];
            const TileUV uv = tile_uv(AT, A, ((T.keylo >> 7) & 1u) == 0, tex, w1, w2, lane);
            const uint32_t baseColor = sample_level(tex, uv.U, uv.V, uv.mip, !uv.minify);
            uint32_t ch2 = 0;
            if (tex.num_layers >= 2) {
                const uint32_t nmr = sample_level(tex, uv.U, uv.V, uv.mip, !uv.minify, 1);
                if (match) {
                    // the fragment's triangle: meshlet and primitive from its rank
                    const uint32_t rank = ~T.keylo;
                    const uint32_t prim = ((rank >> 8) & 7u) * 16u + ((rank >> 3) & 15u);
                    const uint32_t mi = V.meshlet_offset + ((rank >> 11) - V.seq_base);
                    const uint8_t* I = V.m.indices + (size_t)mi * 384;
                    float nv[3][3], tv[3][3];
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const uint32_t nt = V.m.normal_tangents[(size_t)mi * 64 + (I[q * 128 + prim] & 63u)];
                        unpack_normal(AT, nt, nv[q]);
                        unpack_normal(AT, nt >> 16, tv[q]);
                    }
                    const uint32_t handed = (uint32_t)((V.m.tangent_handedness[mi] >> (I[prim] & 63u)) & 1u) << 31;
                    float N[3], Tw[3];
                    lerp_to_world(AT, nv, uv.b, V.o2w, N);
                    lerp_to_world(AT, tv, uv.b, V.o2w, Tw);
                    normal_map(AT, N, Tw, handed, nmr);
                    ch2 = gbuffer_pack(AT, N, nmr);
                }
            }
            if (match) {
                V.aux[off] = baseColor;
                V.aux[(size_t)V.width * V.height + off] = ch2;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------- binning

constexpr uint32_t kSortedBinningPairs = 128;

// Ascending sort of one u32 per lane across the wave (bitonic network).
__device__ __forceinline__ uint32_t wave_sort_u32(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v, (int)j);
            const bool up = (lane & k) == 0;       // this lane's block sorts ascending
            const bool low = (lane & j) == 0;      // lower element of the pair
            v = (low == up) ? min(v, o) : max(v, o);
        }
    }
    return v;
}

// Screen binning of the setup records (the role of DistributeToBins and
// BinQueue::InsertBin, Rasterizer.cpp:663-776, with fixed-capacity bins):
// each wave expands a chunk of records into their (record, bin) pairs with
// an exclusive scan and inserts the pairs in parallel, so a triangle
// spanning many bins costs one pass of the wave, not a serial loop.
// Bins of the rectangle a meshlet's records cover, up to this many: their
// pair counts, then their next free slots, live in LDS.
constexpr uint32_t kRectBins = 1024;

struct BinScratch {                     // one wave's slice of the block's LDS
    // per record: first pair, first bin, bin-box width | bins per row << 16,
    // record index, overflow flag (64 records a call in bin_records, 128 in
    // bin_meshlet)
    uint32_t start[128], org[128], bw[128], ridx[128], spill[128];
    uint32_t rect[128];                 // rect bin of the first bin | box width << 16
    uint32_t marks[64];                 // record starting at each pair of a round, + 1
    uint32_t slots[kRectBins];          // per rect bin: pairs, then next slot
};

// Inserts the wave's records (one per lane where `valid`) into the screen
// bins of their views.
__device__ void bin_records(const DrawParams& p, bool valid, const TriRecord& rec, uint32_t ridx, uint32_t lane,
                            BinScratch& sm, unsigned long long& entries) {
    uint32_t nb = 0, org = 0, bw = 0, pitch = 0;
    if (valid) {
        const ViewParams& V = view_of(p, rec.flags >> 16);
        uint32_t bmin, bmax;
        render_bbox(rec.pos[0], rec.pos[1], rec.pos[2], V.width / 2, V.height / 2, bmin, bmax);
        uint32_t bx0 = (uint32_t)lo16(bmin) >> p.bin_shift, by0 = (uint32_t)hi16(bmin) >> p.bin_shift;
        uint32_t bx1 = (uint32_t)(lo16(bmax) - 1) >> p.bin_shift, by1 = (uint32_t)(hi16(bmax) - 1) >> p.bin_shift;
        bw = bx1 - bx0 + 1;
        nb = bw * (by1 - by0 + 1);
        // first bin of the record's box, in the global bin numbering
        org = V.bin_base + by0 * V.bins_x + bx0;
        pitch = V.bins_x;
    }
    const uint32_t incl = wave_scan_add(nb);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    entries += total;
    __builtin_amdgcn_wave_barrier();   // the previous call's reads of `sm` are done
    // Starts must not decrease with the lane for the search below: an
    // invalid lane (nb = 0) gets the empty range at the running total. The
    // valid lanes of a clipped triangle's fan need not be contiguous (setup
    // culls some fan triangles), so a sentinel start here would hide the
    // lanes after it.
    sm.start[lane] = incl - nb;
    sm.org[lane] = org;
    sm.bw[lane] = bw | (pitch << 16);
    sm.ridx[lane] = ridx;
    sm.spill[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t j = base + lane;
        // (bin << 6 | record lane) of this lane's pair; the sentinel sorts last
        uint32_t item = 0xFFFFFFFFu;
        if (j < total) {
            uint32_t k = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1) {
                if (sm.start[k + step] <= j) k += step;
            }
            const uint32_t q = j - sm.start[k];
            const uint32_t w = sm.bw[k] & 0xFFFFu, pitch_k = sm.bw[k] >> 16;
            const uint32_t dy = q / w, dx = q - dy * w;
            item = ((sm.org[k] + dy * pitch_k + dx) << 6) | k;
        }
        // Small triangles of one meshlet share few bins: sort the wave's
        // pairs by bin so each run of equal bins takes its slots with one
        // atomic. Chunks of large triangles (many pairs) have mostly
        // distinct bins and skip the sort. `total` is wave-uniform.
        if (total <= kSortedBinningPairs) item = wave_sort_u32(item, lane);
        const uint32_t bin = item >> 6;
        const bool ok = item != 0xFFFFFFFFu;
        const uint32_t prev = __shfl_up(item, 1);
        const bool head = ok && (lane == 0 || (prev >> 6) != bin);
        const unsigned long long heads = __ballot(head);
        const uint32_t nvalid = (uint32_t)__popcll(__ballot(ok));
        const unsigned long long later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const uint32_t runEnd = later ? (uint32_t)__ffsll((long long)later) - 1 : nvalid;
        uint32_t got = 0;
        if (head) got = atomicAdd(&p.bin_counts[bin], runEnd - lane);
        // first lane of this lane's run: highest head at or below it
        const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
        const uint32_t runStart = upto ? 63u - (uint32_t)__clzll((long long)upto) : 0u;
        const uint32_t slot = __shfl(got, (int)runStart) + (lane - runStart);
        if (ok) {
            const uint32_t k = item & 63u;
            if (slot < p.bin_capacity) p.bin_lists[(size_t)bin * p.bin_capacity + slot] = sm.ridx[k];
            else sm.spill[k] = 1;
        }
    }
    // A record with a pair past its bin's capacity goes to the overflow list
    // once and is flagged, so the bins that did take it skip it: the overflow
    // raster draws it over its whole render box.
    __builtin_amdgcn_wave_barrier();
    const bool spilled = valid && sm.spill[lane] != 0;
    const unsigned long long spm = __ballot(spilled);
    if (spm) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&p.scratch->num_spill, (uint32_t)__popcll(spm));
        base = __shfl(base, 0);
        if (spilled) {
            p.spill[base + mbcnt(spm)] = ridx;
            p.records[ridx].flags = rec.flags | kRecSpilled;
        }
    }
}

// Per-half minimum of two packed u16 values.
__device__ __forceinline__ uint32_t min_u16x2(uint32_t a, uint32_t b) {
    return min(a & 0xFFFFu, b & 0xFFFFu) | (min(a >> 16, b >> 16) << 16);
}

// Bins a lane walks for its own records; a wave with a larger record box
// walks its pairs in parallel instead.
constexpr uint32_t kLanePairs = 16;

// Inserts up to 128 records of one view (record k = lane + 64 h in slot h of
// each lane) into the screen bins with one device atomic per bin they
// reach. The records' pairs are counted per bin of the rectangle their bin
// boxes span, in LDS; each bin with pairs takes its run of slots with one
// atomicAdd; a second walk over the pairs writes them to those slots. Small
// triangles of a meshlet share few bins, and a bin counter that every
// overlapping wave increments once per pair run makes each later atomic on
// it wait for the earlier ones. Rectangles of more than kRectBins bins go
// through bin_records, 64 records at a time.
__device__ void bin_meshlet(const DrawParams& p, const bool (&valid)[2], const TriRecord (&rec)[2],
                            const uint32_t (&ridx)[2], uint32_t lane, BinScratch& sm,
                            unsigned long long& entries) {
    const unsigned long long vm0 = __ballot(valid[0]), vm1 = __ballot(valid[1]);
    if (!(vm0 | vm1)) return;
    const int src = __ffsll((long long)(vm0 ? vm0 : vm1)) - 1;
    const uint32_t flags = (uint32_t)__builtin_amdgcn_readlane((int)(vm0 ? rec[0].flags : rec[1].flags), src);
    const ViewParams& V = view_of(p, flags >> 16);
    // bin boxes, x | y << 16; the rectangle's corners (the high corner as the
    // minimum of the complement) and the largest box, by wave reductions
    uint32_t lo[2], hi[2], nb[2], rlo = 0xFFFFFFFFu, rhi = 0xFFFFFFFFu, big = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        lo[h] = hi[h] = nb[h] = 0;
        if (valid[h]) {
            uint32_t bmin, bmax;
            render_bbox(rec[h].pos[0], rec[h].pos[1], rec[h].pos[2], V.width / 2, V.height / 2, bmin, bmax);
            lo[h] = ((uint32_t)lo16(bmin) >> p.bin_shift) | (((uint32_t)hi16(bmin) >> p.bin_shift) << 16);
            hi[h] = ((uint32_t)(lo16(bmax) - 1) >> p.bin_shift) | (((uint32_t)(hi16(bmax) - 1) >> p.bin_shift) << 16);
            nb[h] = ((hi[h] & 0xFFFFu) - (lo[h] & 0xFFFFu) + 1) * ((hi[h] >> 16) - (lo[h] >> 16) + 1);
            rlo = min_u16x2(rlo, lo[h]);
            rhi = min_u16x2(rhi, ~hi[h]);
            big = max(big, nb[h]);
        }
    }
    rlo = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan(rlo, 0xFFFFFFFFu, min_u16x2), 63);
    rhi = ~(uint32_t)__builtin_amdgcn_readlane((int)wave_scan(rhi, 0xFFFFFFFFu, min_u16x2), 63);
    // with every record's box at most kLanePairs bins each lane walks its
    // own records' pairs (`steps` steps); otherwise the wave walks all pairs
    // in parallel, 64 at a time
    const uint32_t maxnb = wave_max(big);
    const bool wide = maxnb > kLanePairs;
    const uint32_t steps = wide ? 0u : maxnb;
    const uint32_t rx0 = rlo & 0xFFFFu, ry0 = rlo >> 16;
    const uint32_t rw = (rhi & 0xFFFFu) - rx0 + 1, rh = (rhi >> 16) - ry0 + 1;
    if (rw * rh > kRectBins) {
        if (vm0) bin_records(p, valid[0], rec[0], ridx[0], lane, sm, entries);
        if (vm1) bin_records(p, valid[1], rec[1], ridx[1], lane, sm, entries);
        return;
    }
    const uint32_t R = rw * rh;
    if (counting(p)) entries += (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_add(nb[0] + nb[1]), 63);
    uint32_t bw[2], org[2], rect[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint32_t bx0 = lo[h] & 0xFFFFu, by0 = lo[h] >> 16;
        bw[h] = valid[h] ? (hi[h] & 0xFFFFu) - bx0 + 1 : 0u;
        org[h] = V.bin_base + by0 * V.bins_x + bx0;                  // first bin, global numbering
        rect[h] = (by0 - ry0) * rw + (bx0 - rx0);                    // first bin, rectangle numbering
    }
    __builtin_amdgcn_wave_barrier();   // earlier reads of `sm` are done
    for (uint32_t r = lane; r < R; r += 64) sm.slots[r] = 0;
    uint32_t total = 0, st[2] = { 0, 0 };
    if (wide) {
        // the pairs in record order: record k holds pairs
        // [start[k], start[k] + nb[k])
        const uint32_t i0 = wave_scan_add(nb[0]), i1 = wave_scan_add(nb[1]);
        const uint32_t total0 = (uint32_t)__builtin_amdgcn_readlane((int)i0, 63);
        total = total0 + (uint32_t)__builtin_amdgcn_readlane((int)i1, 63);
        st[0] = i0 - nb[0];
        st[1] = total0 + i1 - nb[1];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            sm.start[h * 64 + lane] = st[h];
            sm.org[h * 64 + lane] = org[h];
            sm.rect[h * 64 + lane] = rect[h] | (bw[h] << 16);
            sm.ridx[h * 64 + lane] = ridx[h];
            sm.spill[h * 64 + lane] = 0;
        }
    }
    // Record owning pair j0 + lane of the round of 64 pairs from j0: each
    // record starting in the round marks its first pair, a max-scan carries
    // the latest mark to the pairs after it, and pairs before the round's
    // first mark belong to the record current at the previous round's end.
    auto owner = [&](uint32_t j0, uint32_t& carry) -> uint32_t {
        __builtin_amdgcn_wave_barrier();
        sm.marks[lane] = 0;
#pragma unroll
        for (int h = 0; h < 2; h++)
            if (nb[h] && st[h] - j0 < 64u) sm.marks[st[h] - j0] = h * 64 + lane + 1;
        __builtin_amdgcn_wave_barrier();
        const uint32_t m = wave_scan(sm.marks[lane], 0u, [](uint32_t a, uint32_t b) { return max(a, b); });
        const uint32_t own = m ? m : carry;
        carry = (uint32_t)__builtin_amdgcn_readlane((int)own, 63);
        return own - 1;
    };
    __builtin_amdgcn_wave_barrier();
    // pass 1: pairs per rectangle bin
    {
        uint32_t r[2] = { rect[0], rect[1] }, dx[2] = { 0, 0 };
        for (uint32_t t = 0; t < steps; t++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (t < nb[h]) {
                    atomicAdd(&sm.slots[r[h]], 1u);
                    r[h]++;
                    if (++dx[h] == bw[h]) { dx[h] = 0; r[h] += rw - bw[h]; }
                }
            }
        }
    }
    if (wide) {
        uint32_t carry = 0;
        for (uint32_t j0 = 0; j0 < total; j0 += 64) {
            const uint32_t k = owner(j0, carry), j = j0 + lane;
            if (j < total) {
                const uint32_t q = j - sm.start[k], rc = sm.rect[k], w = rc >> 16;
                const uint32_t dy = q / w, dx = q - dy * w;
                atomicAdd(&sm.slots[(rc & 0xFFFFu) + dy * rw + dx], 1u);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    // one device atomic per bin with pairs: the bin's first slot (four per
    // lane issued before the first result is waited for)
    for (uint32_t g0 = 0; g0 < R; g0 += 256) {
        uint32_t first_slot[4];
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t r = g0 + i * 64 + lane;
            first_slot[i] = 0;
            if (g0 + i * 64 < R) {
                const uint32_t n = r < R ? sm.slots[r] : 0u;
                if (n) {
                    const uint32_t dy = r / rw, dx = r - dy * rw;
                    first_slot[i] = atomicAdd(&p.bin_counts[V.bin_base + (ry0 + dy) * V.bins_x + rx0 + dx], n);
                }
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < 4; i++)
            if (g0 + i * 64 + lane < R) sm.slots[g0 + i * 64 + lane] = first_slot[i];
    }
    __builtin_amdgcn_wave_barrier();
    // pass 2: each pair takes the next slot of its bin
    bool spill[2] = { false, false };
    {
        uint32_t r[2] = { rect[0], rect[1] }, bin[2] = { org[0], org[1] }, dx[2] = { 0, 0 };
        for (uint32_t t = 0; t < steps; t++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (t < nb[h]) {
                    const uint32_t slot = atomicAdd(&sm.slots[r[h]], 1u);
                    if (slot < p.bin_capacity) p.bin_lists[(size_t)bin[h] * p.bin_capacity + slot] = ridx[h];
                    else spill[h] = true;
                    r[h]++;
                    bin[h]++;
                    if (++dx[h] == bw[h]) {
                        dx[h] = 0;
                        r[h] += rw - bw[h];
                        bin[h] += V.bins_x - bw[h];
                    }
                }
            }
        }
    }
    if (wide) {
        uint32_t carry = 0;
        for (uint32_t j0 = 0; j0 < total; j0 += 64) {
            const uint32_t k = owner(j0, carry), j = j0 + lane;
            if (j < total) {
                const uint32_t q = j - sm.start[k], rc = sm.rect[k], w = rc >> 16;
                const uint32_t dy = q / w, dx = q - dy * w;
                const uint32_t slot = atomicAdd(&sm.slots[(rc & 0xFFFFu) + dy * rw + dx], 1u);
                const uint32_t bin = sm.org[k] + dy * V.bins_x + dx;
                if (slot < p.bin_capacity) p.bin_lists[(size_t)bin * p.bin_capacity + slot] = sm.ridx[k];
                else sm.spill[k] = 1;
            }
        }
        __builtin_amdgcn_wave_barrier();
        spill[0] = sm.spill[lane] != 0;
        spill[1] = sm.spill[64 + lane] != 0;
    }
    // A record with a pair past its bin's capacity goes to the overflow list
    // once and is flagged, so the bins that did take it skip it.
    const bool s0 = valid[0] && spill[0], s1 = valid[1] && spill[1];
    const unsigned long long m0 = __ballot(s0), m1 = __ballot(s1);
    if (m0 | m1) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&p.scratch->num_spill, (uint32_t)(__popcll(m0) + __popcll(m1)));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
        if (s0) {
            p.spill[base + mbcnt(m0)] = ridx[0];
            p.records[ridx[0]].flags = rec[0].flags | kRecSpilled;
        }
        if (s1) {
            p.spill[base + (uint32_t)__popcll(m0) + mbcnt(m1)] = ridx[1];
            p.records[ridx[1]].flags = rec[1].flags | kRecSpilled;
        }
    }
}

// Binning of the meshlet records, then clipping of the queued triangles
// (one wave each; the fan triangles go straight into the bins from the
// wave's registers).
template <bool Alpha>
__global__ void __launch_bounds__(256) k_bin_scatter(DrawParams p) {
    __shared__ BinScratch s_bins[4];
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * 4;
    const uint32_t first = blockIdx.x * 4 + wid;
    unsigned long long entries = 0;
    // The first meshlet's record count loads beside the visible count
    // instead of after it: every meshlet has a count slot, and a wave past
    // the visible count ignores what it read.
    uint32_t i = first, cnt = 0;
    if (i < p.total_meshlets) cnt = p.meshlet_counts[i];
    const uint32_t nvis = p.scratch->num_visible;
    const uint32_t nclip = min(p.scratch->num_clip, p.clip_capacity);
    const uint32_t q0 = waves - 1 - first;
    // one visible meshlet's records at a time, the next one's count in flight
    while (i < nvis) {
        const uint32_t next = i + waves;
        TriRecord r0 = {}, r1 = {};
        if (lane < cnt) r0 = p.records[(size_t)i * 128 + lane];
        if (lane + 64 < cnt) r1 = p.records[(size_t)i * 128 + 64 + lane];
        const uint32_t ncnt = next < nvis ? p.meshlet_counts[next] : 0u;
        const bool valid[2] = { lane < cnt, lane + 64 < cnt };
        const TriRecord rec[2] = { r0, r1 };
        const uint32_t ridx[2] = { i * 128 + lane, i * 128 + 64 + lane };
        bin_meshlet(p, valid, rec, ridx, lane, s_bins[wid], entries);
        i = next;
        cnt = ncnt;
    }
    // Queued triangles go to the last waves of the grid, which have no
    // meshlet when the grid holds more waves than visible meshlets: clipping
    // then runs beside the binning instead of after it.
    for (uint32_t q = q0; q < nclip; q += waves) {
        TriRecord rec;
        uint32_t ridx = 0;
        const bool valid = clip_one<Alpha, true>(p, clip_src(p.clip_queue[q]), lane, rec, ridx, q * 8);
        const bool v2[2] = { valid, false };
        const TriRecord r2[2] = { rec, rec };
        const uint32_t x2[2] = { ridx, 0 };
        bin_meshlet(p, v2, r2, x2, lane, s_bins[wid], entries);
    }
    if (counting(p) && lane == 0 && entries) atomicAdd(&p.counters->bin_entries, entries);
}

// One block per bin of any view. The bin's keys live in LDS in the same 4x4
// tile order as the vis-buffer. The block loads up to kThreads records at a
// time and clips their render boxes to the bin; for bins of up to 32 pixels
// each record's thread marks the 4x4 tiles of its box the triangle can reach
// in a 64-bit mask, and a block-wide exclusive scan of the masks' popcounts
// numbers the chunk's candidate tiles. The numbering is cut into equal
// ranges, one per 16-lane group; a group walks its range tile by tile (lane =
// pixel), moving to the next record as the range crosses it, and keeps the
// nearest fragment with an LDS 64-bit max (64-pixel bins walk every tile of
// each box instead). The finished bin is
// written to HBM once, one contiguous 4-pixel-high strip of tiles at a
// time. After a clear the bin starts from the clear key, otherwise from the
// vis-buffer's current contents.
//
// One block per bin, in bin order. (Keeping neighbouring bins on one XCD's
// L2 by remapping block b to the (b/8)-th bin of the (b%8)-th eighth of the
// range measured 4% slower: consecutive blocks then run far-apart bins, and
// the records they share were already in the MALL.)
template <int BS, int kThreads, bool Alpha>
__device__ void raster_bin(const DrawParams& p, uint32_t bin, unsigned long long (&keys)[BS * BS],
                           BinTri<BS> (&tl)[kThreads], unsigned long long (&tmasks)[bin_masks(BS, kThreads)],
                           uint32_t (&s_wave_total)[kThreads / kWave],
                           uint32_t (&talpha)[Alpha ? kThreads : 1]) {
    constexpr int kWaves = kThreads / kWave;
    constexpr int kGroups = kThreads / 16;
    const uint32_t vidx = view_of_bin(p, bin);
    const ViewParams& V = view_of(p, vidx);
    const uint32_t lbin = bin - V.bin_base;
    const uint32_t bx = lbin % V.bins_x, by = lbin / V.bins_x;
    const uint32_t W = (uint32_t)V.width, H = (uint32_t)V.height;
    const uint32_t x0 = bx * BS, y0 = by * BS;
    const uint32_t x1 = min(x0 + BS, W), y1 = min(y0 + BS, H);
    const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    // The first kPrefetchEntries list entries are loaded together with the
    // count (entries past the count are never used; the read stays inside
    // the bin's list), which takes one dependent load off most bins; the
    // rest wait for the count.
    constexpr uint32_t kPrefetchEntries = kThreads < 128 ? kThreads : 128;
    const uint32_t* list = p.bin_lists + (size_t)bin * p.bin_capacity;
    const uint32_t first_entry = tid < min(p.bin_capacity, kPrefetchEntries) ? list[tid] : 0u;
    const uint32_t count = p.bin_counts[bin];
    const uint32_t flagged = p.bin_flags[bin];
    const uint32_t n = min(count, p.bin_capacity);
    // records on the overflow list and the fan triangles of triangles that
    // found the clip queue full were drawn into the view's spill keys, which
    // flagged the bins they reached
    const bool merge = flagged != 0;
    // the visible count has had its last reader (the scatter): leave it
    // cleared for the next draw's cull
    if (bin == 0 && tid == 0) p.scratch->num_visible = 0;
    if (tid == 0 && flagged) p.bin_flags[bin] = 0u;
    if (n == 0 && !merge && !V.fresh) return;  // uniform across the block
    const int hw = V.width / 2, hh = V.height / 2;
    // view fields the tile loop needs, in registers: the compiler does not
    // move a load of the view table (global memory for batches) across the
    // LDS atomics, so a field read inside the loop is a load and a wait per
    // tile
    const float clearDepth = V.clear_depth;
    const unsigned long long clearKey = (unsigned long long)fbits(clearDepth) << 32;
    const uint32_t rowElems = (x1 - x0) * 4;   // elements of one 4-row strip of tiles
    unsigned long long* gkeys = V.keys;

    for (uint32_t i = tid; i < BS * BS; i += kThreads) {
        uint32_t strip = i / (BS * 4), rem = i % (BS * 4);
        unsigned long long val = clearKey;
        if (!V.fresh && rem < rowElems && y0 + strip * 4 < y1)
            val = gkeys[(size_t)(y0 + strip * 4) * W + x0 * 4 + rem];
        keys[i] = val;
    }

    const bool countFrags = counting(p);
    uint32_t covered = 0, passed = 0, tiles = 0;
    const uint32_t g = tid >> 4, px = tid & 3, py = (tid >> 2) & 3;
    constexpr bool kMasked = BS <= 32;
    constexpr uint32_t kTilesX = BS / 4;                                   // tiles per bin row
    constexpr uint32_t kTileShift = kTilesX == 2 ? 1u : (kTilesX == 4 ? 2u : 3u);
    for (uint32_t base = 0; base < n; base += kThreads) {
        const uint32_t e = base + tid;
        const uint32_t cnt = min(n - base, (uint32_t)kThreads);          // records of this chunk
        uint32_t ntiles = 0;
        unsigned long long tmask = 0;      // candidate tiles of this thread's record
        BinTri<BS> t;
        if (e < n) {
            const uint32_t r = (base == 0 && e < kPrefetchEntries) ? first_entry : list[e];
            const TriRecord rec = p.records[r];
            uint32_t bmin, bmax;
            TriLds f;
            load_tri(rec, hw, hh, f, bmin, bmax);
            const uint32_t cx0 = max((uint32_t)lo16(bmin), x0), cy0 = max((uint32_t)hi16(bmin), y0);
            const uint32_t cx1 = min((uint32_t)lo16(bmax), x1), cy1 = min((uint32_t)hi16(bmax), y1);
            if constexpr (Alpha) talpha[tid] = (rec.flags & kRecAlpha) ? r : 0xFFFFFFFFu;
            t.a12 = f.a12; t.a20 = f.a20; t.a01 = f.a01;
            t.b12 = f.b12; t.b20 = f.b20; t.b01 = f.b01;
            t.z0 = f.z0; t.z10 = f.z10; t.z20 = f.z20;
            t.keylo = f.keylo;
            t.w0 = t.w1 = t.w2 = 0;
            t.li0 = 0;
            uint32_t tw = 0, th = 0;
            // a record on the overflow list is drawn by the overflow raster
            if (cx1 > cx0 && cy1 > cy0 && !(rec.flags & kRecSpilled)) {
                tw = (cx1 - cx0) >> 2;
                th = (cy1 - cy0) >> 2;
                // edge values at pixel (cx0, cy0); sums wrap in 32 bits as a
                // direct evaluation does (DrawTriangle, Rasterizer.h:268-296)
                t.w0 = edge_at(f.e0, f.a12, f.b12, cx0, cy0);
                t.w1 = edge_at(f.e1, f.a20, f.b20, cx0, cy0);
                t.w2 = edge_at(f.e2, f.a01, f.b01, cx0, cy0);
                // the box starts on a tile corner: its first tile's index in the
                // bin's 4x4-tiled LDS layout
                t.li0 = ((cx0 - x0) << 2) + (cy0 - y0) * BS;
            }
            if constexpr (kMasked) {
                // The tiles of the box that the triangle can cover, one tile
                // row at a time. Each edge's largest value over a tile's 16
                // pixels is V = D + 4A tx + 4B ty, D = e + max(0, 3A) +
                // max(0, 3B) at the box's first tile; a tile where one V is
                // negative has no covered pixel. Along a row V >= 0 holds on
                // one side of q = -(D + 4B ty) / 4A: tx >= q for A > 0, tx <= q
                // for A < 0. q comes from a float reciprocal (relative error a
                // few 2^-24, far below the 2^-10 margin taken outward over the
                // |q| <= 16 that matter), so the mask keeps every tile with a
                // covered pixel. Edge values inside a render box stay below 2^31
                // in magnitude (|e| < 2^27, |A|, |B| < 2^17, pixel coordinates
                // below 2^12), so D + 4B ty is exact.
                unsigned long long mask = 0;
                if (tw) {
                    const int32_t A[3] = {f.a12, f.a20, f.a01}, Bv[3] = {f.b12, f.b20, f.b01};
                    const int32_t W0[3] = {t.w0, t.w1, t.w2};
                    int32_t D[3];
                    float nr[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        D[i] = W0[i] + max(0, 3 * A[i]) + max(0, 3 * Bv[i]);
                        nr[i] = A[i] != 0 ? -__builtin_amdgcn_rcpf((float)(4 * A[i])) : 0.0f;
                    }
                    for (uint32_t ty = 0; ty < th; ty++) {
                        int32_t lo = 0, hi = (int32_t)tw - 1;
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            const int32_t R = D[i] + __mul24(Bv[i], (int32_t)(4 * ty));
                            const float q = fminf(fmaxf((float)R * nr[i], -16.0f), 16.0f);
                            if (A[i] > 0) lo = max(lo, (int32_t)ceilf(q - 0x1p-10f));
                            else if (A[i] < 0) hi = min(hi, (int32_t)floorf(q + 0x1p-10f));
                            else if (R < 0) hi = -1;
                        }
                        if (hi >= lo)
                            mask |= (unsigned long long)((2u << hi) - (1u << lo)) << (ty * kTilesX);
                    }
                }
                tmask = mask;
                ntiles = (uint32_t)__popcll(mask);
            } else {
                ntiles = tw * th;
                t.tw = tw ? tw : 1u;
                t.r0 = (int32_t)(((uint32_t)f.b12 << 2) - (uint32_t)__mul24(f.a12, (int32_t)(4 * (t.tw - 1))));
                t.r1 = (int32_t)(((uint32_t)f.b20 << 2) - (uint32_t)__mul24(f.a20, (int32_t)(4 * (t.tw - 1))));
                t.r2 = (int32_t)(((uint32_t)f.b01 << 2) - (uint32_t)__mul24(f.a01, (int32_t)(4 * (t.tw - 1))));
            }
        }
        const uint32_t incl = wave_scan_add(ntiles);
        if (lane == 63) s_wave_total[wid] = incl;
        __syncthreads();
        uint32_t prev = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < (uint32_t)kWaves; w++) {
            uint32_t val = s_wave_total[w];
            if (w < wid) prev += val;
            total += val;
        }
        t.start = prev + incl - ntiles;
        tl[tid] = t;
        if (e >= n) tl[tid].start = 0xFFFFFFFFu;

        // DrawTriangle's coverage test and depth (Rasterizer.h:281-296) for
        // this lane's pixel of a tile of record T, given its edge values and
        // LDS key index
        BinTri<BS> T;
        uint32_t ai = 0xFFFFFFFFu;
        auto shade = [&](uint32_t w0, uint32_t w1, uint32_t w2, uint32_t li) {
            if constexpr (!Alpha) {
                if ((int32_t)(w0 | w1 | w2) >= 0) {
                    const float depth = __builtin_fmaf((float)(int32_t)w1, T.z10,
                                                       __builtin_fmaf((float)(int32_t)w2, T.z20, T.z0));
                    if (countFrags) covered++;
                    if (depth > clearDepth) {
                        if (countFrags) passed++;
                        // an LDS max without a returned value: no wait on the
                        // old key (a read-then-max would stall every step)
                        const unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
                        atomicMax(&keys[li], key);
                    }
                }
            } else {
                const bool cov = (int32_t)(w0 | w1 | w2) >= 0;
                const float depth = __builtin_fmaf((float)(int32_t)w1, T.z10,
                                                   __builtin_fmaf((float)(int32_t)w2, T.z20, T.z0));
                bool pass = cov && depth > clearDepth;
                if (ai != 0xFFFFFFFFu && (__ballot(pass) & (0xFFFFull << (lane & 48u)))) {
                    const bool ok = alpha_test_tile(p.tables, p.alpha[ai], ((T.keylo >> 7) & 1u) == 0,
                                                    V.textures, (int32_t)w1, (int32_t)w2, lane);
                    pass = pass && ok;
                }
                if (countFrags && cov) covered++;
                if (pass) {
                    if (countFrags) passed++;
                    const unsigned long long key = ((unsigned long long)fbits(depth) << 32) | T.keylo;
                    atomicMax(&keys[li], key);
                }
            }
        };

        if constexpr (kMasked) {
            // A record's candidate tiles are the set bits of its mask, the
            // first at position `start` of the chunk's tile list. Group g
            // takes positions [j, j1): it finds the record holding j (the
            // last one starting at or before it), skips that record's first
            // j - start tiles and walks set bits, moving on to the next
            // record with tiles when a mask runs out. (Laying the tiles out
            // as an item list first cost a serial loop per record, as long
            // as the block's largest triangle, and a barrier.)
            tmasks[tid] = tmask;
            __syncthreads();
            const uint32_t per = (total + kGroups - 1) / kGroups;
            uint32_t j = min(g * per, total);
            const uint32_t j1 = min(j + per, total);
            if (countFrags && j < j1) tiles += j1 - j;
            if (j < j1) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = kThreads / 2; step > 0; step >>= 1)
                    if (tl[k + step].start <= j) k += step;
                T = tl[k];
                if constexpr (Alpha) ai = talpha[k];
                unsigned long long m = tmasks[k];
                for (uint32_t skip = j - T.start; skip; skip--) m &= m - 1ull;
                for (;;) {
                    const uint32_t b = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1ull;
                    const uint32_t tx = b & (kTilesX - 1), ty = b >> kTileShift;
                    const int32_t mx = (int32_t)(px + 4 * tx), my = (int32_t)(py + 4 * ty);
                    shade((uint32_t)T.w0 + (uint32_t)__mul24(T.a12, mx) + (uint32_t)__mul24(T.b12, my),
                          (uint32_t)T.w1 + (uint32_t)__mul24(T.a20, mx) + (uint32_t)__mul24(T.b20, my),
                          (uint32_t)T.w2 + (uint32_t)__mul24(T.a01, mx) + (uint32_t)__mul24(T.b01, my),
                          T.li0 + (tx << 4) + ty * (4 * BS) + px + (py << 2));
                    if (++j >= j1) break;
                    if (!m) {          // tiles remain in the range: a later record has them
                        do {
                            m = tmasks[++k];
                        } while (!m);
                        T = tl[k];
                        if constexpr (Alpha) ai = talpha[k];
                    }
                }
            }
            __syncthreads();
        } else {
            __syncthreads();
            // group g walks tiles [j, j1) of the block's tile list
            const uint32_t per = (total + kGroups - 1) / kGroups;
            uint32_t j = min(g * per, total);
            const uint32_t j1 = min(j + per, total);
            if (countFrags && j < j1) tiles += j1 - j;
            if (j < j1) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = kThreads / 2; step > 0; step >>= 1) {
                    if (tl[k + step].start <= j) k += step;
                }
                T = tl[k];
                if constexpr (Alpha) ai = talpha[k];
                uint32_t next = (k + 1 < (uint32_t)kThreads) ? tl[k + 1].start : 0xFFFFFFFFu;
                const uint32_t pt = j - T.start;
                const uint32_t ty0 = pt / T.tw;
                uint32_t tx = pt - ty0 * T.tw;
                // Edge values and the LDS key index of this lane's pixel, stepped
                // from tile to tile (DrawTriangle's stepping, Rasterizer.h:268-
                // 296): +4A per tile in x, and from the last tile of a row to the
                // first of the next the record's row step. Sums wrap in 32 bits,
                // as a direct evaluation does.
                uint32_t w0, w1, w2, s0, s1, s2, li, liRow;
                auto enter = [&](uint32_t ty) {
                    const int32_t mx = (int32_t)(px + 4 * tx), my = (int32_t)(py + 4 * ty);
                    w0 = (uint32_t)T.w0 + (uint32_t)__mul24(T.a12, mx) + (uint32_t)__mul24(T.b12, my);
                    w1 = (uint32_t)T.w1 + (uint32_t)__mul24(T.a20, mx) + (uint32_t)__mul24(T.b20, my);
                    w2 = (uint32_t)T.w2 + (uint32_t)__mul24(T.a01, mx) + (uint32_t)__mul24(T.b01, my);
                    s0 = (uint32_t)T.a12 << 2;
                    s1 = (uint32_t)T.a20 << 2;
                    s2 = (uint32_t)T.a01 << 2;
                    li = T.li0 + (tx << 4) + ty * (4 * BS) + px + (py << 2);
                    liRow = 4 * BS - 16 * (T.tw - 1);
                };
                enter(ty0);
                for (; j < j1; j++) {
                    if (j >= next) {      // step to the record owning tile j
                        do {
                            k++;
                            next = (k + 1 < (uint32_t)kThreads) ? tl[k + 1].start : 0xFFFFFFFFu;
                        } while (j >= next);
                        T = tl[k];
                        if constexpr (Alpha) ai = talpha[k];
                        tx = 0;
                        enter(0);
                    }
                    shade(w0, w1, w2, li);
                    // next tile, without a branch
                    const bool rowEnd = ++tx == T.tw;
                    w0 += rowEnd ? (uint32_t)T.r0 : s0;
                    w1 += rowEnd ? (uint32_t)T.r1 : s1;
                    w2 += rowEnd ? (uint32_t)T.r2 : s2;
                    li += rowEnd ? liRow : 16u;
                    tx = rowEnd ? 0u : tx;
                }
            }
            __syncthreads();
        }
    }
    // each thread writes back the keys it initialised, so the next bin's
    // initialisation needs no extra barrier; an overflowing bin takes the
    // maximum with the spill keys of its pixels and zeroes them for the next
    // draw
    for (uint32_t i = tid; i < BS * BS; i += kThreads) {
        uint32_t strip = i / (BS * 4), rem = i % (BS * 4);
        if (rem < rowElems && y0 + strip * 4 < y1) {
            const size_t gi = (size_t)(y0 + strip * 4) * W + x0 * 4 + rem;
            unsigned long long val = keys[i];
            if (merge) {
                const unsigned long long sk = V.spill_keys[gi];
                if (sk) {
                    val = sk > val ? sk : val;
                    V.spill_keys[gi] = 0ull;
                }
            }
            gkeys[gi] = val;
        }
    }
    // the bin's count, read by every thread before the first barrier of the
    // record loop (count > 0 means the loop ran), is cleared for the next draw
    if (tid == 0 && count) p.bin_counts[bin] = 0;
    if (countFrags) {
        for (int d = 32; d > 0; d >>= 1) {
            covered += __shfl_xor(covered, d);
            passed += __shfl_xor(passed, d);
        }
        if (lane == 0) {
            atomicAdd(&p.counters->fragments_covered, (unsigned long long)covered);
            atomicAdd(&p.counters->fragments_passed, (unsigned long long)passed);
        }
        // one lane of each 16-lane group counted the group's tiles
        if ((tid & 15u) == 0 && tiles) atomicAdd(&p.counters->tile_lanes, 16ull * tiles);
        if (tid == 0 && n) atomicAdd(&p.counters->bins_rasterized, 1ull);
    }
}

// The overflow raster: one wave per record on the overflow list draws it
// over its whole render box into the view's spill keys (raster_now), marking
// the bins it passes a fragment in; the bins' rasters skip the record and
// merge the spill keys. The work is that of the listed records' areas,
// however many bins or views the launch has.
template <bool Alpha>
__device__ void overflow_raster(const DrawParams& p, uint32_t first_wave, uint32_t waves) {
    const uint32_t lane = lane_id();
    const uint32_t ns = min(p.scratch->num_spill, p.spill_capacity);
    for (uint32_t q = first_wave; q < ns; q += waves) {
        const uint32_t r = p.spill[q];
        const TriRecord rec = p.records[r];
        AlphaRecord a = {};
        if constexpr (Alpha) {
            if (rec.flags & kRecAlpha) a = p.alpha[r];
        }
        raster_now<Alpha>(p, rec, a, lane);
    }
}

// Blocks of the bin-order launch that run the overflow raster (they exit at
// once when no pair overflowed, so an empty list costs no launch of its own).
constexpr uint32_t kOverflowBlocks = 256;

// Orders the bins for the raster, most records first: bin b goes to class
// c = bit length of its record count (0 for an empty bin) and takes a slot
// in that class's list. Work per bin grows with its records, and crowded
// bins (at 32-pixel bins the largest hold about four times the mean) would
// otherwise finish last on an idle machine. Blocks past the bins' run the
// overflow raster.
template <bool Alpha>
__global__ void __launch_bounds__(256) k_bin_order(DrawParams p) {
    __shared__ uint32_t s_count[kBinBuckets], s_base[kBinBuckets], s_max;
    const uint32_t tid = threadIdx.x;
    const uint32_t order_blocks = (p.total_bins + 255) / 256;
    if (blockIdx.x >= order_blocks) {
        const uint32_t w = (blockIdx.x - order_blocks) * 4 + tid / kWave;
        clip_overflow<Alpha>(p, w, kOverflowBlocks * 4);
        overflow_raster<Alpha>(p, w, kOverflowBlocks * 4);
        return;
    }
    const uint32_t bin = blockIdx.x * 256 + tid;
    if (tid < kBinBuckets) s_count[tid] = 0;
    if (tid == 0) s_max = 0;
    __syncthreads();
    uint32_t cls = 0, rank = 0, offered = 0;
    const bool valid = bin < p.total_bins;
    if (valid) {
        offered = p.bin_counts[bin];
        const uint32_t n = min(offered, p.bin_capacity);
        cls = n ? min(kBinBuckets - 1, 32u - (uint32_t)__clz(n)) : 0u;
        rank = atomicAdd(&s_count[cls], 1u);
    }
    // the largest bin, for launch-shape choices: one device atomic per block,
    // and none when the block's largest is not above the value already there
    offered = wave_max(offered);
    if (lane_id() == 0 && offered) atomicMax(&s_max, offered);
    __syncthreads();
    if (tid == 0 && s_max > p.scratch->max_bin_count) atomicMax(&p.scratch->max_bin_count, s_max);
    if (tid < kBinBuckets) s_base[tid] = s_count[tid] ? atomicAdd(&p.scratch->bucket[tid], s_count[tid]) : 0u;
    __syncthreads();
    if (valid) p.bin_order[(size_t)cls * p.total_bins + s_base[cls] + rank] = bin;
}

// One block per bin; block i rasterizes the i-th bin of the heavy-first
// order, so the blocks dispatched first (the hardware starts blocks in
// index order) get the crowded bins. (Taking the position from a ticket
// counter instead measured much slower: every block start then waits for one
// contended atomic.)
template <int BS, int kThreads, bool Alpha>
__global__ void __launch_bounds__(kThreads) k_bin_raster(DrawParams p) {
    __shared__ unsigned long long keys[BS * BS];
    __shared__ BinTri<BS> tl[kThreads];
    __shared__ unsigned long long tmasks[bin_masks(BS, kThreads)];
    __shared__ uint32_t s_wave_total[kThreads / kWave];
    __shared__ uint32_t talpha[Alpha ? kThreads : 1];
    __shared__ uint32_t s_bin;
    if (threadIdx.x == 0) {
        // every class size in one load, then the class holding position
        // blockIdx.x without branches (a search that stops at the first
        // class large enough loads the sizes one after another)
        uint32_t n[kBinBuckets];
#pragma unroll
        for (uint32_t c = 0; c < kBinBuckets; c++) n[c] = p.scratch->bucket[c];
        const uint32_t item = blockIdx.x;
        uint32_t cls = 0, start = 0, acc = 0;
        bool found = false;
#pragma unroll
        for (int c = (int)kBinBuckets - 1; c >= 0; c--) {
            const bool here = !found && item - acc < n[c];
            cls = here ? (uint32_t)c : cls;
            start = here ? acc : start;
            found = found || here;
            acc += n[c];
        }
        // (every position is in a class)
        s_bin = found ? p.bin_order[(size_t)cls * p.total_bins + (item - start)] : item;
    }
    __syncthreads();
    raster_bin<BS, kThreads, Alpha>(p, s_bin, keys, tl, tmasks, s_wave_total, talpha);
}

// Clears every scratch counter: at the end of a direct draw (the binned
// path clears num_visible in the bin raster and the others in the cull)
// and in place of the cull of a launch without meshlets.
__global__ void k_reset_counters(DrawScratch* s) {
    s->num_records = 0;
    s->num_clip = 0;
    s->num_visible = 0;
    s->num_spill = 0;
    s->num_clip_overflow = 0;
    s->max_bin_count = 0;
    for (uint32_t c = 0; c < kBinBuckets; c++) s->bucket[c] = 0;
}

__global__ void k_clear(unsigned long long* keys, uint32_t n, unsigned long long key) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) keys[i] = key;
}

// Vis-buffer keys -> reference layers: surface id (MeshletOffset +
// MeshletId) * 128 + PrimId (Shading.cpp:328) and depth.
__global__ void k_decode(const unsigned long long* keys, int width, int height, unsigned long long clear_key,
                         uint32_t clear_color, const DrawEntry* draws, uint32_t num_draws, uint32_t* ids,
                         float* depth, int linear) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = (uint32_t)width * (uint32_t)height;
    if (i >= n) return;
    unsigned long long k = keys[i];
    uint32_t id = clear_color;
    uint32_t db = (uint32_t)(clear_key >> 32);
    if (k != clear_key) {
        uint32_t rank = ~(uint32_t)k;
        db = (uint32_t)(k >> 32);
        uint32_t seq = rank >> 11;
        uint32_t prim = ((rank >> 8) & 7u) * 16u + ((rank >> 3) & 15u);
        uint32_t lo = 0, hi = num_draws;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (draws[mid].seq_base <= seq) lo = mid; else hi = mid;
        }
        uint32_t meshlet = draws[lo].meshlet_offset + (seq - draws[lo].seq_base);
        id = meshlet * 128u + prim;
    }
    uint32_t o = i;
    if (linear) {
        // i is a tiled offset; recover (x, y)
        uint32_t W = (uint32_t)width;
        uint32_t tileRow = i / (W * 4u);
        uint32_t rem = i - tileRow * W * 4u;
        uint32_t tx = rem >> 4, q = rem & 15u;
        uint32_t x = tx * 4 + (q & 3u), y = tileRow * 4 + (q >> 2);
        o = y * W + x;
    }
    if (ids) ids[o] = id;
    if (depth) depth[o] = bitsf(db);
}

__global__ void k_detile32(const uint32_t* src, int width, uint32_t n, uint32_t* out, int linear) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = i;
    if (linear) {
        const uint32_t W = (uint32_t)width;
        const uint32_t tileRow = i / (W * 4u);
        const uint32_t rem = i - tileRow * W * 4u;
        const uint32_t tx = rem >> 4, q = rem & 15u;
        o = (tileRow * 4 + (q >> 2)) * W + tx * 4 + (q & 3u);
    }
    out[o] = src[i];
}

__global__ void k_fill32(uint32_t* buf, uint32_t n, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) buf[i] = value;
}

// Overdraw counts and depth -> the reference's layers: the counter word
// with each 16-bit half the clear colour's half plus the count, saturated
// (FS_Overdraw's _mm512_adds_epu16, Shading.cpp:336-337), and the depth.
__global__ void k_decode_overdraw(const unsigned long long* counts, const uint32_t* depth, uint32_t n,
                                  uint32_t clear_color, int width, uint32_t* words, uint32_t* depth_out, int linear) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long c = counts[i];
    const uint32_t hi = (uint32_t)min((unsigned long long)(clear_color >> 16) + (c >> 32), 0xFFFFull);
    const uint32_t lo = (uint32_t)min((unsigned long long)(clear_color & 0xFFFFu) + (c & 0xFFFFFFFFull), 0xFFFFull);
    uint32_t o = i;
    if (linear) {
        const uint32_t W = (uint32_t)width;
        const uint32_t tileRow = i / (W * 4u);
        const uint32_t rem = i - tileRow * W * 4u;
        const uint32_t tx = rem >> 4, q = rem & 15u;
        o = (tileRow * 4 + (q >> 2)) * W + tx * 4 + (q & 3u);
    }
    if (words) words[o] = (hi << 16) | lo;
    if (depth_out) depth_out[o] = depth[i];
}

// The bin raster for the launch's bin size and workgroup size: 128 or 256
// threads for every bin size, 512 or 1024 for 32- and 64-pixel bins (a
// crowded large bin then has more waves to spread its tiles over).
template <int BS, int T, bool Alpha>
void launch_bin_raster_one(const DrawParams& p, hipStream_t s, Profiler* prof) {
    launch_kernel(prof, "bin_raster", k_bin_raster<BS, T, Alpha>, dim3(p.total_bins), dim3(T), 0, s, p);
}

template <int BS, bool Alpha>
bool launch_bin_raster_bs(const DrawParams& p, hipStream_t s, Profiler* prof) {
    switch (p.raster_threads) {
    case 128: launch_bin_raster_one<BS, 128, Alpha>(p, s, prof); return true;
    case 256: launch_bin_raster_one<BS, 256, Alpha>(p, s, prof); return true;
    case 512:
        if constexpr (BS >= 32) {
            launch_bin_raster_one<BS, 512, Alpha>(p, s, prof);
            return true;
        }
        return false;
    case 1024:
        if constexpr (BS >= 32) {
            launch_bin_raster_one<BS, 1024, Alpha>(p, s, prof);
            return true;
        }
        return false;
    default: return false;
    }
}


template <bool Alpha>
bool launch_bin_raster(const DrawParams& p, hipStream_t s, Profiler* prof) {
    switch (p.bin_size) {
    case 8: return launch_bin_raster_bs<8, Alpha>(p, s, prof);
    case 16: return launch_bin_raster_bs<16, Alpha>(p, s, prof);
    case 32: return launch_bin_raster_bs<32, Alpha>(p, s, prof);
    case 64: return launch_bin_raster_bs<64, Alpha>(p, s, prof);
    default: return false;
    }
}

void launch_front(const DrawParams& p, hipStream_t s, Profiler* prof, bool separate_clip) {
    if (!p.total_meshlets) hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, s, p.scratch);
    if (p.total_meshlets) {     // a batch of empty views still clears through the bin raster
        if (p.total_meshlets <= 256u * 1024u)
            launch_kernel(prof, "cull", k_cull<1>, dim3((p.total_meshlets + 255) / 256), dim3(256), 0, s, p);
        else
            launch_kernel(prof, "cull", k_cull<4>, dim3((p.total_meshlets + 1023) / 1024), dim3(256), 0, s, p);
    }
    if (p.total_meshlets) {
        const uint32_t want = (p.total_meshlets + 3) / 4;
        if (p.alpha)
            launch_kernel(prof, "mesh_setup", k_mesh_setup<true>,
                          dim3(min(want, resident_blocks<k_mesh_setup<true>, 256>())), dim3(256), 0, s, p);
        else
            launch_kernel(prof, "mesh_setup", k_mesh_setup<false>,
                          dim3(min(want, resident_blocks<k_mesh_setup<false>, 256>())), dim3(256), 0, s, p);
    }
    if (separate_clip) {
        if (p.alpha) launch_kernel(prof, "clip", k_clip<true>, dim3(256), dim3(256), 0, s, p);
        else launch_kernel(prof, "clip", k_clip<false>, dim3(256), dim3(256), 0, s, p);
    }
}

}  // namespace

void warm_launch_shapes() {
    (void)resident_blocks<k_mesh_setup<false>, 256>();
    (void)resident_blocks<k_mesh_setup<true>, 256>();
    (void)resident_blocks<k_raster_direct<false>, 256>();
    (void)resident_blocks<k_raster_direct<true>, 256>();
    (void)resident_blocks<k_raster_overdraw, 256>();
    (void)resident_blocks<k_gbuffer, 256>();
    (void)resident_blocks<k_bin_scatter<false>, 256>();
    (void)resident_blocks<k_bin_scatter<true>, 256>();
}

hipError_t launch_reset_counters(DrawScratch* scratch, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, s, scratch);
    return hipGetLastError();
}

hipError_t launch_clear(unsigned long long* keys, uint32_t n, unsigned long long key, hipStream_t s) {
    uint32_t blocks = min((n + 255) / 256, 8192u);
    hipLaunchKernelGGL(k_clear, dim3(blocks), dim3(256), 0, s, keys, n, key);
    return hipGetLastError();
}

hipError_t launch_fill32(uint32_t* buf, uint32_t n, uint32_t value, hipStream_t s) {
    uint32_t blocks = min((n + 255) / 256, 8192u);
    hipLaunchKernelGGL(k_fill32, dim3(blocks), dim3(256), 0, s, buf, n, value);
    return hipGetLastError();
}

hipError_t launch_decode_overdraw(const unsigned long long* counts, const uint32_t* depth, uint32_t n,
                                  uint32_t clear_color, int width, uint32_t* words, uint32_t* depth_out, int linear,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_decode_overdraw, dim3((n + 255) / 256), dim3(256), 0, s, counts, depth, n, clear_color,
                       width, words, depth_out, linear);
    return hipGetLastError();
}

hipError_t launch_draw_overdraw(const DrawParams& p, hipStream_t s, Profiler* prof) {
    launch_front(p, s, prof, true);
    {
        const uint32_t want = max((2 * p.total_meshlets + 3) / 4 + 16, 64u);
        launch_kernel(prof, "raster_overdraw", k_raster_overdraw,
                      dim3(min(want, resident_blocks<k_raster_overdraw, 256>())), dim3(256), 0, s, p);
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, s, p.scratch);
    return hipGetLastError();
}

hipError_t launch_draw_deferred(const DrawParams& p, hipStream_t s, Profiler* prof) {
    launch_front(p, s, prof, true);
    const uint32_t want = max((2 * p.total_meshlets + 3) / 4 + 16, 64u);
    launch_kernel(prof, "raster_direct", k_raster_direct<true>,
                  dim3(min(want, resident_blocks<k_raster_direct<true>, 256>())), dim3(256), 0, s, p);
    launch_kernel(prof, "gbuffer", k_gbuffer, dim3(min(want, resident_blocks<k_gbuffer, 256>())), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, s, p.scratch);
    return hipGetLastError();
}

hipError_t launch_detile32(const uint32_t* src, int width, uint32_t n, uint32_t* out, int linear, hipStream_t s) {
    hipLaunchKernelGGL(k_detile32, dim3((n + 255) / 256), dim3(256), 0, s, src, width, n, out, linear);
    return hipGetLastError();
}

hipError_t launch_decode(const unsigned long long* keys, int width, int height, unsigned long long clear_key,
                         uint32_t clear_color, const DrawEntry* draws, uint32_t num_draws, uint32_t* ids,
                         float* depth, int linear, hipStream_t s) {
    uint32_t n = (uint32_t)width * (uint32_t)height;
    hipLaunchKernelGGL(k_decode, dim3((n + 255) / 256), dim3(256), 0, s, keys, width, height, clear_key, clear_color,
                       draws, num_draws, ids, depth, linear);
    return hipGetLastError();
}

hipError_t launch_cull(const DevMeshlets& m, uint32_t first, uint32_t count, const float* planes,
                       const Occlusion* occ, const ApproxTables& tables, uint16_t* bitmap, uint32_t* visible,
                       hipStream_t s) {
    Planes5 pl;
    for (int i = 0; i < 20; i++) pl.p[i] = planes[i];
    Occlusion o = {};
    if (occ) o = *occ;
    hipLaunchKernelGGL(k_cull_range, dim3((count + 255) / 256), dim3(256), 0, s, m, first, count, pl, o, tables,
                       bitmap, visible);
    return hipGetLastError();
}

hipError_t launch_hiz_build(const unsigned long long* keys, const Occlusion& occ, float* depth_map, hipStream_t s) {
    for (uint32_t lv = 0; lv + 1 < occ.mip_levels; lv++) {
        const uint32_t w = occ.level_width[lv];
        const uint32_t h = ((occ.height - 1) >> (lv + 1)) + 1;
        const uint32_t n = w * h;
        hipLaunchKernelGGL(k_hiz_level, dim3((n + 255) / 256), dim3(256), 0, s, keys, occ.width, occ.height, lv,
                           lv ? depth_map + occ.level_offset[lv - 1] : nullptr, lv ? occ.level_width[lv - 1] : 0u,
                           depth_map + occ.level_offset[lv], w, h);
    }
    return hipGetLastError();
}

hipError_t launch_draw_direct(const DrawParams& p, hipStream_t s, Profiler* prof) {
    launch_front(p, s, prof, true);
    {
        const uint32_t want = max((2 * p.total_meshlets + 3) / 4 + 16, 64u);
        if (p.alpha)
            launch_kernel(prof, "raster_direct", k_raster_direct<true>,
                          dim3(min(want, resident_blocks<k_raster_direct<true>, 256>())), dim3(256), 0, s, p);
        else
            launch_kernel(prof, "raster_direct", k_raster_direct<false>,
                          dim3(min(want, resident_blocks<k_raster_direct<false>, 256>())), dim3(256), 0, s, p);
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, s, p.scratch);
    return hipGetLastError();
}

hipError_t launch_draw_binned(const DrawParams& p, hipStream_t s, Profiler* prof) {
    launch_front(p, s, prof, false);   // the scatter kernel clips
    {
        // about one chunk per wave: two per meshlet plus room for clipped records
        const uint32_t want = max((2 * p.total_meshlets + 3) / 4 + 16, 64u);
        if (p.alpha)
            launch_kernel(prof, "bin_scatter", k_bin_scatter<true>,
                          dim3(min(want, resident_blocks<k_bin_scatter<true>, 256>())), dim3(256), 0, s, p);
        else
            launch_kernel(prof, "bin_scatter", k_bin_scatter<false>,
                          dim3(min(want, resident_blocks<k_bin_scatter<false>, 256>())), dim3(256), 0, s, p);
    }
    {
        const dim3 grid((p.total_bins + 255) / 256 + kOverflowBlocks);
        if (p.alpha) launch_kernel(prof, "bin_order", k_bin_order<true>, grid, dim3(256), 0, s, p);
        else launch_kernel(prof, "bin_order", k_bin_order<false>, grid, dim3(256), 0, s, p);
    }
    // records of alpha-tested materials need the variant with the alpha test
    if (p.alpha) {
        if (!launch_bin_raster<true>(p, s, prof)) return hipErrorInvalidValue;
    } else {
        if (!launch_bin_raster<false>(p, s, prof)) return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace glimp
