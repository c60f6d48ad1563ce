// Differentiable texture-sampling resolve for gfx950.
//
// Forward: per covered pixel, the surface from the vis-buffer key, the
// triangle's barycentrics (IntersectTriangle's formula with correctly
// rounded divisions; coverage and barycentrics are constants of the pass),
// then
//   U,V = uv_0 + b1 (uv_1 - uv_0) + b2 (uv_2 - uv_0)
//   col = color_0 + b1 (color_1 - color_0) + b2 (color_2 - color_0)
//   out = Sample(U, V) * col
// with fp32 texel mip chains. Sample keeps the resolve sampler's choices
// (Texture.h:412-459, SurfaceSampler of Shading.cpp:6-10): mip level from the
// UV gradients, one filter per 4x4 tile (any minified pixel -> nearest texel
// of level clamp(mip), else bilinear at level 0), Repeat wrap; the bilinear
// weights are continuous (texel centres at +0.5, clamped at the edges) so
// that the output is differentiable in U and V.
//
// Backward: one pass over the pixels writes, per pixel, four texel records
// (texel index, weight * dL/dsample) and three vertex records (vertex index,
// barycentric weight * (dL/dcol, dL/dU, dL/dV)) into fixed slots. The
// records are sorted by index with a stable radix sort, and every run of
// equal indices is summed in slot order: in 256-record tiles by one thread
// per run piece, then across tiles in tile order. The order of every
// floating-point addition depends only on the inputs, so the gradients are
// bitwise reproducible; no floating-point atomics are used.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "diff.h"
#include "surface.h"

namespace glimp {
namespace {

constexpr uint32_t kTile = 256;

struct PixelState {
    bool active;
    uint32_t mi, vi[3];
    float b1, b2, U, V;
    int32_t tex, level;
    bool linear;
    float fu, fv;                  // U, V minus their wrap period
    int32_t tx, ty;
    bool cx, cy, ex, ey;
    float xs, ys;
};

__device__ __forceinline__ const DiffView& view_at(const DiffParams& p, uint32_t v) {
    return p.views ? p.views[v] : p.view;
}

// View owning global pixel slot t.
__device__ __forceinline__ uint32_t view_of_pixel(const DiffParams& p, uint32_t t) {
    uint32_t lo = 0, hi = p.num_views;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (view_at(p, mid).pixel_base <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Pixel slot t of the global numbering: its view, and the pixel's row-major
// index and 4x4-tiled index inside that view (slots run over the tiles).
struct PixelSlot {
    uint32_t view, x, y, tiled;
    bool inside;
};

__device__ __forceinline__ PixelSlot pixel_slot(const DiffParams& p, uint32_t t) {
    PixelSlot s;
    s.inside = t < p.total_pixels;
    s.view = s.inside ? view_of_pixel(p, t) : 0;
    const DiffView& V = view_at(p, s.view);
    s.tiled = t - (s.inside ? V.pixel_base : t);
    const uint32_t W = (uint32_t)V.width;
    const uint32_t tile = s.tiled >> 4, q = s.tiled & 15u, tilesX = W / 4;
    uint32_t tx;
    const uint32_t ty = div_small(tile, tilesX, tx);
    s.x = tx * 4 + (q & 3u);
    s.y = ty * 4 + (q >> 2);
    return s;
}

// Surface, barycentrics and sampler choices of pixel slot t (tiled order:
// 16 lanes = one 4x4 tile of one view). Every lane of the wave must call it.
__device__ void diff_pixel(const DiffParams& p, const PixelSlot& ps, PixelState& S) {
    const DiffView& VW = view_at(p, ps.view);
    const bool inside = ps.inside;
    const uint32_t x = ps.x, y = ps.y;
    const uint32_t lane = __lane_id();
    const unsigned long long groupMask = 0xFFFFull << (lane & 48u);
    const ApproxTables& T = p.tables;

    const unsigned long long key = inside ? VW.keys[ps.tiled] : VW.clear_key;
    const bool active = inside && !(bitsf((uint32_t)(key >> 32)) <= 0.0f);
    S.active = active;
    S.tex = -1;
    S.level = 0;
    S.linear = false;
    S.b1 = S.b2 = S.U = S.V = 0.0f;
    float gradU[2] = { 0, 0 }, gradV[2] = { 0, 0 };
    uint32_t material = 0xFFFFFFFFu;
    if (active) {
        uint32_t tri;
        decode_surface(key, VW.draws, VW.num_draws, S.mi, tri);
        const uint8_t* I = p.m.indices + (size_t)S.mi * 384;
        const float* P = p.m.positions + (size_t)S.mi * 192;
        float c[3][4], uv[3][2];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t vi = I[k * 128 + tri];
            S.vi[k] = vi;
            const float px = P[vi], py = P[64 + vi], pz = P[128 + vi];
#pragma unroll
            for (int j = 0; j < 4; j++) c[k][j] = xform(VW.object_to_clip, j, px, py, pz);
            const float2 t2 = p.uv[(size_t)S.mi * 64 + vi];
            uv[k][0] = t2.x;
            uv[k][1] = t2.y;
        }
        material = p.m.material_id[S.mi];
        const float su = __builtin_fmaf((float)(int32_t)x, VW.scale_x, VW.center_x);
        const float sv = __builtin_fmaf((float)(int32_t)y, VW.scale_y, VW.center_y);
        float bary[3], ddx[3], ddy[3];
        intersect_triangle<true>(T, c, su, sv, VW.kx, VW.ky, bary, ddx, ddy);
        S.b1 = bary[1];
        S.b2 = bary[2];
        const float t10x = uv[1][0] - uv[0][0], t10y = uv[1][1] - uv[0][1];
        const float t20x = uv[2][0] - uv[0][0], t20y = uv[2][1] - uv[0][1];
        S.U = __builtin_fmaf(bary[2], t20x, __builtin_fmaf(bary[1], t10x, uv[0][0]));
        S.V = __builtin_fmaf(bary[2], t20y, __builtin_fmaf(bary[1], t10y, uv[0][1]));
        gradU[0] = __builtin_fmaf(t20x, ddx[2], t10x * ddx[1]);
        gradV[0] = __builtin_fmaf(t20y, ddx[2], t10y * ddx[1]);
        gradU[1] = __builtin_fmaf(t20x, ddy[2], t10x * ddy[1]);
        gradV[1] = __builtin_fmaf(t20y, ddy[2], t10y * ddy[1]);
    }

    // material waterfall per 4x4 tile with the tile-wide filter vote
    bool done = !active || material == 0xFFFFFFFFu || material >= p.num_materials;
    for (;;) {
        const unsigned long long todo = __ballot(!done);
        if (!todo) break;
        const unsigned long long mine = todo & groupMask;
        const uint32_t leader = mine ? (uint32_t)__ffsll((long long)mine) - 1 : lane;
        const uint32_t lmat = __shfl(material, (int)leader);
        int32_t texIdx = -1;
        if (mine) texIdx = p.materials[lmat].texture;
        if (texIdx >= (int32_t)p.num_textures) texIdx = -1;
        int32_t mip = -64;
        uint32_t TW = 1, TH = 1, TL = 1;
        if (texIdx >= 0) {
            const DiffTex& D = p.textures[texIdx];
            TW = D.width;
            TH = D.height;
            TL = D.levels;
            const float su2 = (float)TW * (float)TW, sv2 = (float)TH * (float)TH;
            const float dx = __builtin_fmaf(gradU[0], gradU[0], gradV[0] * gradV[0]) * su2;
            const float dy = __builtin_fmaf(gradU[1], gradU[1], gradV[1] * gradV[1]) * sv2;
            const float mx = (dx > dy) ? dx : dy;
            mip = (int32_t)(fbits(mx) - (127u << 23)) >> 24;
        }
        const unsigned long long vote = __ballot(active && mip > 0);
        const bool minify = (vote & groupMask) != 0;
        if (!done && material == lmat) {
            done = true;
            if (texIdx >= 0) {
                S.tex = texIdx;
                S.linear = !minify;
                const float floorU = floorf(S.U), floorV = floorf(S.V);
                S.fu = S.U - floorU;
                S.fv = S.V - floorV;
                if (S.linear) {
                    S.level = 0;
                    S.xs = S.fu * (float)TW - 0.5f;
                    S.ys = S.fv * (float)TH - 0.5f;
                    S.cx = S.xs > 0.0f;
                    S.cy = S.ys > 0.0f;
                    const float xf = S.cx ? S.xs : 0.0f, yf = S.cy ? S.ys : 0.0f;
                    S.tx = (int32_t)xf;
                    S.ty = (int32_t)yf;
                    S.ex = S.tx < (int32_t)TW - 1;
                    S.ey = S.ty < (int32_t)TH - 1;
                    if (!S.ex) S.tx = (int32_t)TW - 1;
                    if (!S.ey) S.ty = (int32_t)TH - 1;
                } else {
                    const int32_t m = min(max(mip, 0), (int32_t)TL - 1);
                    S.level = m;
                    const int32_t Wm = (int32_t)(TW >> m), Hm = (int32_t)(TH >> m);
                    S.tx = min((int32_t)(S.fu * (float)Wm), Wm - 1);
                    S.ty = min((int32_t)(S.fv * (float)Hm), Hm - 1);
                }
            }
        }
    }
}

// Sample value, its texels and weights, and its U/V derivatives.
struct SampleEval {
    float4 val, du, dv;
    uint64_t texel[4];
    float w[4];
    int n;
};

__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }
__device__ __forceinline__ float4 operator-(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 operator*(float s, float4 a) { return make_float4(s * a.x, s * a.y, s * a.z, s * a.w); }
__device__ __forceinline__ float4 operator*(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ __forceinline__ float4 fma4(float s, float4 a, float4 b) {
    return make_float4(__builtin_fmaf(s, a.x, b.x), __builtin_fmaf(s, a.y, b.y), __builtin_fmaf(s, a.z, b.z),
                       __builtin_fmaf(s, a.w, b.w));
}
__device__ __forceinline__ float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float dot4(float4 a, float4 b) {
    return __builtin_fmaf(a.x, b.x, __builtin_fmaf(a.y, b.y, __builtin_fmaf(a.z, b.z, a.w * b.w)));
}

__device__ void sample_eval(const DiffParams& p, const PixelState& S, SampleEval& E) {
    E.n = 0;
    E.du = E.dv = f4(0.0f);
    if (S.tex < 0) {
        E.val = f4(1.0f);
        return;
    }
    const DiffTex& D = p.textures[S.tex];
    const uint64_t base = D.level_offset[S.level];
    if (!S.linear) {
        const uint64_t Wm = D.width >> S.level;
        E.n = 1;
        E.texel[0] = base + (uint64_t)S.ty * Wm + (uint64_t)S.tx;
        E.w[0] = 1.0f;
        E.val = p.texels[E.texel[0]];
        return;
    }
    const float xf = S.cx ? S.xs : 0.0f, yf = S.cy ? S.ys : 0.0f;
    const float fx = S.ex ? xf - (float)S.tx : 0.0f, fy = S.ey ? yf - (float)S.ty : 0.0f;
    const uint64_t x0 = (uint64_t)S.tx, y0 = (uint64_t)S.ty;
    const uint64_t x1 = S.ex ? x0 + 1 : x0, y1 = S.ey ? y0 + 1 : y0;
    E.n = 4;
    E.texel[0] = base + y0 * D.width + x0;
    E.texel[1] = base + y0 * D.width + x1;
    E.texel[2] = base + y1 * D.width + x0;
    E.texel[3] = base + y1 * D.width + x1;
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    E.w[0] = gx * gy;
    E.w[1] = fx * gy;
    E.w[2] = gx * fy;
    E.w[3] = fx * fy;
    const float4 t00 = p.texels[E.texel[0]], t10 = p.texels[E.texel[1]];
    const float4 t01 = p.texels[E.texel[2]], t11 = p.texels[E.texel[3]];
    const float4 d0 = t10 - t00, d1 = t11 - t01;
    const float4 l0 = fma4(fx, d0, t00), l1 = fma4(fx, d1, t01);
    E.val = fma4(fy, l1 - l0, l0);
    const float sx = (S.cx && S.ex) ? (float)D.width : 0.0f, sy = (S.cy && S.ey) ? (float)D.height : 0.0f;
    E.du = sx * add4(gy * d0, fy * d1);
    E.dv = sy * (l1 - l0);
}

__device__ __forceinline__ float4 vertex_color(const DiffParams& p, const PixelState& S) {
    const float4* C = p.color + (size_t)S.mi * 64;
    const float4 c0 = C[S.vi[0]], c1 = C[S.vi[1]], c2 = C[S.vi[2]];
    return fma4(S.b2, c2 - c0, fma4(S.b1, c1 - c0, c0));
}

__device__ __forceinline__ size_t out_index(const DiffParams& p, const PixelSlot& ps) {
    const DiffView& V = view_at(p, ps.view);
    return (size_t)V.pixel_base + (size_t)ps.y * (uint32_t)V.width + ps.x;
}

__global__ void __launch_bounds__(256) k_diff_forward(DiffParams p) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const PixelSlot ps = pixel_slot(p, t);
    PixelState S;
    diff_pixel(p, ps, S);
    if (!ps.inside) return;
    float4 o = f4(0.0f);
    if (S.active) {
        SampleEval E;
        sample_eval(p, S, E);
        o = E.val * vertex_color(p, S);
    }
    p.out[out_index(p, ps)] = o;
}

// Per-pixel gradient records (fixed slots 4t.. and 3t..).
__global__ void __launch_bounds__(256) k_diff_records(DiffParams p) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const PixelSlot ps = pixel_slot(p, t);
    PixelState S;
    diff_pixel(p, ps, S);
    if (!ps.inside) return;
    uint32_t tkeys[4] = { p.tex_sentinel, p.tex_sentinel, p.tex_sentinel, p.tex_sentinel };
    float4 tvals[4] = { f4(0.0f), f4(0.0f), f4(0.0f), f4(0.0f) };
    uint32_t vkeys[3] = { p.vtx_sentinel, p.vtx_sentinel, p.vtx_sentinel };
    float4 vcol[3] = { f4(0.0f), f4(0.0f), f4(0.0f) }, vuv[3] = { f4(0.0f), f4(0.0f), f4(0.0f) };
    if (S.active) {
        SampleEval E;
        sample_eval(p, S, E);
        const float4 g = p.grad_out[out_index(p, ps)];
        const float4 col = vertex_color(p, S);
        const float4 gval = g * col;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < E.n) {
                tkeys[k] = (uint32_t)E.texel[k];
                tvals[k] = E.w[k] * gval;
            }
        }
        const float gU = dot4(gval, E.du), gV = dot4(gval, E.dv);
        const float4 gcol = g * E.val;
        const float w[3] = { (1.0f - S.b1) - S.b2, S.b1, S.b2 };
#pragma unroll
        for (int k = 0; k < 3; k++) {
            vkeys[k] = S.mi * 64 + S.vi[k];
            vcol[k] = w[k] * gcol;
            vuv[k] = make_float4(w[k] * gU, w[k] * gV, 0.0f, 0.0f);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        p.tex_keys[(size_t)t * 4 + k] = tkeys[k];
        p.tex_vals[(size_t)t * 4 + k] = tvals[k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p.vtx_keys[(size_t)t * 3 + k] = vkeys[k];
        p.vtx_vals[((size_t)t * 3 + k) * 2 + 0] = vcol[k];
        p.vtx_vals[((size_t)t * 3 + k) * 2 + 1] = vuv[k];
    }
}

// Gradient outputs of a segmented reduction: K float4 per record.
struct TexelOut {
    float4* grad;
    __device__ void write(uint32_t key, const float4* v) const { grad[key] = v[0]; }
};
struct VertexOut {
    float4* color;
    float2* uv;
    __device__ void write(uint32_t key, const float4* v) const {
        color[key] = v[0];
        uv[key] = make_float2(v[1].x, v[1].y);
    }
};

template <int K, class Out>
__global__ void __launch_bounds__(kTile) k_seg_reduce(const uint32_t* keys, const uint32_t* idx, const float4* vals,
                                                      uint32_t n, uint32_t sentinel, Out out, float4* part_first,
                                                      float4* part_last, uint32_t* key_last, uint32_t* spans) {
    __shared__ uint32_t skey[kTile];
    __shared__ float4 sval[kTile][K];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const uint32_t i0 = tile * kTile;
    const uint32_t cnt = min(kTile, n - i0);
    if (tid < cnt) {
        skey[tid] = keys[i0 + tid];
        const uint32_t r = idx[i0 + tid];
#pragma unroll
        for (int k = 0; k < K; k++) sval[tid][k] = vals[(size_t)r * K + k];
    }
    __syncthreads();
    if (tid >= cnt) return;
    const uint32_t key = skey[tid];
    if (key == sentinel) return;
    const uint32_t prev = tid ? skey[tid - 1] : (i0 ? keys[i0 - 1] : sentinel);
    if (tid && prev == key) return;   // not the first record of a piece
    float4 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = sval[tid][k];
    uint32_t j = tid + 1;
    for (; j < cnt && skey[j] == key; j++) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = add4(acc[k], sval[j][k]);
    }
    const bool starts_here = !(tid == 0 && prev == key);
    const bool cont_next = j == cnt && i0 + cnt < n && keys[i0 + cnt] == key;
    if (starts_here && !cont_next) {
        out.write(key, acc);
    } else if (starts_here) {
#pragma unroll
        for (int k = 0; k < K; k++) part_last[(size_t)tile * K + k] = acc[k];
        key_last[tile] = key;
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) part_first[(size_t)tile * K + k] = acc[k];
        spans[tile] = cont_next ? 1u : 0u;
    }
}

// Runs that cross tiles: the tile where a run starts adds the pieces of the
// following tiles in tile order.
template <int K, class Out>
__global__ void k_seg_fixup(uint32_t tiles, uint32_t sentinel, Out out, const float4* part_first,
                            const float4* part_last, const uint32_t* key_last, const uint32_t* spans) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    const uint32_t key = key_last[t];
    if (key == sentinel) return;
    float4 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = part_last[(size_t)t * K + k];
    for (uint32_t u = t + 1; u < tiles; u++) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = add4(acc[k], part_first[(size_t)u * K + k]);
        if (!spans[u]) break;
    }
    out.write(key, acc);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_f32(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

uint32_t bits_for(uint64_t v) {   // bits needed to hold v
    uint32_t b = 0;
    while (b < 32 && (v >> b) != 0) b++;
    return b;
}

template <class T>
hipError_t grow(T*& p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (!count) return hipSuccess;
    return hipMalloc((void**)&p, count * sizeof(T));
}

hipError_t fill_zero(void* p, size_t floats, hipStream_t s) {
    if (!floats) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<size_t>((floats + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, s, (float*)p, floats);
    return hipGetLastError();
}

template <int K, class Out>
hipError_t reduce_records(DiffWorkspace& w, const uint32_t* keys, const float4* vals, uint32_t n, uint32_t sentinel,
                          Out out, hipStream_t s) {
    size_t bytes = w.sort_temp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(w.sort_temp, bytes, keys, w.sorted_keys,
                                             rocprim::counting_iterator<uint32_t>(0), w.sorted_idx, n, 0,
                                             bits_for(sentinel), s);
    if (e != hipSuccess) return e;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>((tiles + 255) / 256, 4096u)), dim3(256), 0, s, w.key_last,
                       (size_t)tiles, sentinel);
    hipLaunchKernelGGL((k_seg_reduce<K, Out>), dim3(tiles), dim3(kTile), 0, s, w.sorted_keys, w.sorted_idx, vals, n,
                       sentinel, out, w.part_first, w.part_last, w.key_last, w.spans);
    hipLaunchKernelGGL((k_seg_fixup<K, Out>), dim3((tiles + 255) / 256), dim3(256), 0, s, tiles, sentinel, out,
                       w.part_first, w.part_last, w.key_last, w.spans);
    return hipGetLastError();
}

}  // namespace

hipError_t diff_workspace_reserve(DiffWorkspace& w, size_t pixels) {
    if (pixels <= w.pixels) return hipSuccess;
    const size_t nt = pixels * 4, nv = pixels * 3, tiles = (nt + kTile - 1) / kTile;
    hipError_t e = grow(w.tex_keys, nt);
    if (e == hipSuccess) e = grow(w.tex_vals, nt);
    if (e == hipSuccess) e = grow(w.vtx_keys, nv);
    if (e == hipSuccess) e = grow(w.vtx_vals, nv * 2);
    if (e == hipSuccess) e = grow(w.sorted_keys, nt);
    if (e == hipSuccess) e = grow(w.sorted_idx, nt);
    if (e == hipSuccess) e = grow(w.part_first, tiles * 2);
    if (e == hipSuccess) e = grow(w.part_last, tiles * 2);
    if (e == hipSuccess) e = grow(w.key_last, tiles);
    if (e == hipSuccess) e = grow(w.spans, tiles);
    if (e != hipSuccess) return e;
    size_t bytes = 0;
    e = rocprim::radix_sort_pairs(nullptr, bytes, w.tex_keys, w.sorted_keys, rocprim::counting_iterator<uint32_t>(0),
                                  w.sorted_idx, (uint32_t)nt, 0, 32, (hipStream_t)0);
    if (e != hipSuccess) return e;
    if (w.sort_temp) (void)hipFree(w.sort_temp);
    w.sort_temp = nullptr;
    e = hipMalloc(&w.sort_temp, bytes);
    if (e != hipSuccess) return e;
    w.sort_temp_bytes = bytes;
    w.pixels = pixels;
    return hipSuccess;
}

void diff_workspace_release(DiffWorkspace& w) {
    void* ptrs[] = { w.tex_keys, w.tex_vals, w.vtx_keys, w.vtx_vals, w.sorted_keys, w.sorted_idx,
                     w.part_first, w.part_last, w.key_last, w.spans, w.sort_temp };
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    w = DiffWorkspace();
}

hipError_t launch_diff_forward(const DiffParams& p, hipStream_t s, Profiler* prof) {
    const uint32_t n = p.total_pixels;
    launch_kernel(prof, "diff_forward", k_diff_forward, dim3((n + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_diff_backward(DiffParams p, DiffWorkspace& w, uint64_t num_texels, uint32_t num_vertices,
                                float4* grad_texels, float2* grad_uv, float4* grad_color, hipStream_t s,
                                Profiler* prof) {
    const uint32_t n = p.total_pixels;
    p.tex_keys = w.tex_keys;
    p.tex_vals = w.tex_vals;
    p.vtx_keys = w.vtx_keys;
    p.vtx_vals = w.vtx_vals;
    p.tex_sentinel = (uint32_t)num_texels;
    p.vtx_sentinel = num_vertices;
    hipError_t e;
    {
        Profiler::Scope sc(prof, "diff_zero_grads", s);
        e = fill_zero(grad_texels, num_texels * 4, s);
        if (e == hipSuccess) e = fill_zero(grad_uv, (size_t)num_vertices * 2, s);
        if (e == hipSuccess) e = fill_zero(grad_color, (size_t)num_vertices * 4, s);
        if (e != hipSuccess) return e;
    }
    launch_kernel(prof, "diff_records", k_diff_records, dim3((n + 255) / 256), dim3(256), 0, s, p);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    {
        Profiler::Scope sc(prof, "diff_reduce_texels", s);
        e = reduce_records<1>(w, w.tex_keys, w.tex_vals, n * 4, p.tex_sentinel, TexelOut{ grad_texels }, s);
        if (e != hipSuccess) return e;
    }
    {
        Profiler::Scope sc(prof, "diff_reduce_vertices", s);
        e = reduce_records<2>(w, w.vtx_keys, w.vtx_vals, n * 3, p.vtx_sentinel, VertexOut{ grad_color, grad_uv }, s);
    }
    return e;
}

}  // namespace glimp
