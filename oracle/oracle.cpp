// CPU oracle: scalar restatement of the reference's single-threaded,
// non-binned draw path (Rasterizer::DrawMeshletsST, Rasterizer.cpp:153-254)
// and of its vis-buffer resolve (Shading.cpp:417-579, 658-773).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Built with -ffp-contract=off and
// without fast-math, so every rounding below is the one written: where the
// reference's -ffast-math build contracts a product into an FMA or turns a
// vector division into VRCP14PS plus a Newton step, this file spells that
// out, as established by oracle/probes/fastmath_probe.cpp.
//
// The reference evaluates 16 SIMD lanes at a time; lanes are independent
// except where noted (packet order decides draw order; the texture filter
// choice is taken over a 4x4 tile). This restatement keeps the 16-lane packet
// structure so draw order and those tile-wide decisions are reproduced.
#include "oracle.h"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- primitives

uint16_t g_rcp[65536];
uint16_t g_rsq[2][32768];

inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

// VRCP14PS (SIMD.h:295), table model; exhaustive agreement with the
// instruction is checked by glimpsw_amd/data/gen_approx14.c --verify.
float rcp14(float xf) {
    uint32_t x = fbits(xf);
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) return bitsf(m ? (x | 0x400000u) : s);
    if (e == 0) {
        if (m == 0) return bitsf(s | 0x7F800000u);
        while (!(m & 0x800000u)) { m <<= 1; e--; }
        m &= 0x7FFFFFu;
        e += 1;
    }
    int re;
    uint32_t mant;
    if (m == 0) { re = 254 - e; mant = 0; }
    else { re = 253 - e; mant = (uint32_t)g_rcp[m >> 7] << 7; }
    if (re >= 255) return bitsf(s | 0x7F800000u);
    if (re <= 0) {
        int sh = 1 - re;
        if (sh > 24) return bitsf(s);
        return bitsf(s | ((mant | 0x800000u) >> sh));
    }
    return bitsf(s | ((uint32_t)re << 23) | mant);
}

// VRSQRT14PS (SIMD.h:294), table model.
float rsqrt14(float xf) {
    uint32_t x = fbits(xf);
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) { if (m) return bitsf(x | 0x400000u); return bitsf(s ? 0xFFC00000u : 0); }
    if (e == 0) {
        if (m == 0) return bitsf(s | 0x7F800000u);
        while (!(m & 0x800000u)) { m <<= 1; e--; }
        m &= 0x7FFFFFu;
        e += 1;
    }
    if (s) return bitsf(0xFFC00000u);
    int ue = e - 127;
    int oc = ue & 1;
    int half = (ue - oc) / 2;
    if (m == 0 && oc == 0) return bitsf((uint32_t)(127 - half) << 23);
    return bitsf(((uint32_t)(126 - half) << 23) | ((uint32_t)g_rsq[oc][m >> 8] << 7));
}

// Vector a/b under -ffast-math on AVX-512 (probe_div, probe_recip):
//   r = rcp14(b); q = a*r; e = fma(q, b, -a); result = fma(-r, e, q)
// For a == 1 the product a*r is exact, giving the 1/w sequence.
float fdiv(float a, float b) {
    float r = rcp14(b);
    float q = a * r;
    float e = fmaf(q, b, -a);
    return fmaf(-r, e, q);
}
float frcp(float b) { return fdiv(1.0f, b); }

// cvtps2dq: round to nearest even, "integer indefinite" when out of range.
int32_t cvt_rne(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)nearbyintf(x);
}

inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

// Packed signed 16-bit pairs (x in the low half, y in the high half).
inline int16_t lo16(uint32_t v) { return (int16_t)(v & 0xFFFF); }
inline int16_t hi16(uint32_t v) { return (int16_t)(v >> 16); }
inline uint32_t pk16(int lo, int hi) { return (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16); }

// IEEE half -> float (vcvtph2ps is exact).
float half_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h & 0x8000) << 16;
    uint32_t e = (h >> 10) & 31, m = h & 1023;
    if (e == 0) {
        if (m == 0) return bitsf(s);
        float f = (float)m * (1.0f / 16777216.0f);  // m * 2^-24, exact
        return s ? -f : f;
    }
    if (e == 31) return bitsf(s | 0x7F800000u | (m << 13));
    return bitsf(s | ((e + 112) << 23) | (m << 13));
}

// simd::mul(mat4, v) with v.w = 1 (SIMD.h:457-464).
void transform(const float* m, float x, float y, float z, float out[4]) {
    for (int c = 0; c < 4; c++) {
        out[c] = fmaf(x, m[0 * 4 + c], fmaf(y, m[1 * 4 + c], fmaf(z, m[2 * 4 + c], 1.0f * m[3 * 4 + c])));
    }
}

constexpr uint32_t kMaxRenderSize = 2896;  // Rasterizer.h:203

// ---------------------------------------------------------------- mesh shader

struct ShadedMeshlet {  // Rasterizer.h:82-99
    uint32_t PrimCount = 0;
    int CullMode = 1;  // 0 None, 1 FrontCCW, 2 FrontCW
    int FragmentShaderId = 0;
    uint8_t Indices[3][128];
    float Position[4][64];
};

// ShadeMeshlet (Shading.cpp:281-307).
void shade_meshlet(const orc_shading_ctx& ctx, uint32_t index, ShadedMeshlet& out) {
    if (ctx.cull_bitmap) {
        uint16_t mask = ctx.cull_bitmap[index / 16];
        if (((mask >> (index % 16)) & 1) == 0) { out.PrimCount = 0; return; }
    }
    const orc_meshlets& ms = *ctx.meshlets;
    uint32_t mi = ctx.meshlet_offset + index;
    const float* pos = ms.positions + (size_t)mi * 192;
    uint32_t nv = ms.num_vertices[mi];
    for (uint32_t i = 0; i < nv; i += 16) {
        for (uint32_t l = 0; l < 16; l++) {
            float c[4];
            transform(ctx.object_to_clip, pos[0 * 64 + i + l], pos[1 * 64 + i + l], pos[2 * 64 + i + l], c);
            for (int k = 0; k < 4; k++) out.Position[k][i + l] = c[k];
        }
    }
    out.PrimCount = ms.num_triangles[mi];
    memcpy(out.Indices, ms.indices + (size_t)mi * 384, 384);
    uint32_t mat = ms.material_id[mi];
    if (mat != UINT32_MAX && ctx.materials) {
        const orc_material& mt = ctx.materials[mat];
        out.CullMode = mt.double_sided ? 0 : 1;
        out.FragmentShaderId = mt.alpha_cutoff < 255 ? 1 : 0;
    }
}

// ---------------------------------------------------------------- packets

struct V4 { float x, y, z, w; };

struct ClipCodes {  // Rasterizer.cpp:51-55
    uint16_t AcceptMask = 0, NonTrivialMask = 0;
    uint8_t OutCodes[16] = {};
};

// Clipper::ComputeClipCodes (Rasterizer.cpp:353-397). The all-inside fast
// path returns the same per-lane masks as the general path, so only the
// general path is stated.
ClipCodes compute_clip_codes(const V4 v[3][16], float bx, float by) {
    ClipCodes cc;
    for (int l = 0; l < 16; l++) {
        uint8_t partial = 0, combined = 255;
        bool trivial = true;
        for (int k = 0; k < 3; k++) {
            const V4& p = v[k][l];
            uint8_t oc = 0;
            if (p.x < -p.w) oc |= 1 << 0;
            if (p.x > +p.w) oc |= 1 << 1;
            if (p.y < -p.w) oc |= 1 << 2;
            if (p.y > +p.w) oc |= 1 << 3;
            if (p.z < -p.w) oc |= 1 << 4;
            if (p.z > +p.w) oc |= 1 << 5;
            partial |= oc;
            combined &= oc;
            trivial &= (fabsf(p.x) < p.w * bx) && (fabsf(p.y) < p.w * by);
        }
        trivial &= (partial & 0x30) == 0;
        bool accept = combined == 0;
        if (accept && trivial) cc.AcceptMask |= (uint16_t)(1u << l);
        if (accept && !trivial) cc.NonTrivialMask |= (uint16_t)(1u << l);
        cc.OutCodes[l] = partial;
    }
    return cc;
}

struct TrianglePacket {  // Rasterizer.h:145-161
    uint32_t Pos0[16], Pos1[16], Pos2[16];
    float Z0[16], Z1[16], Z2[16];
    float W0[16], W1[16], W2[16];
};

// GetBoundingBox + GetRenderBoundingBox (Rasterizer.cpp:331-351).
void render_bbox(const TrianglePacket& t, int l, int hw, int hh, uint32_t& bmin, uint32_t& bmax) {
    uint32_t p0 = t.Pos0[l], p1 = t.Pos1[l], p2 = t.Pos2[l];
    int mnx = std::min(std::min(lo16(p0), lo16(p1)), lo16(p2));
    int mny = std::min(std::min(hi16(p0), hi16(p1)), hi16(p2));
    int mxx = std::max(std::max(lo16(p0), lo16(p1)), lo16(p2));
    int mxy = std::max(std::max(hi16(p0), hi16(p1)), hi16(p2));
    // 32-bit add of 0x0007'0007 to the packed pair: the low half may carry
    // into the high half; then an arithmetic >> 4 per 16-bit half.
    uint32_t mn = pk16(mnx, mny) + 0x00070007u;
    uint32_t mx = pk16(mxx, mxy) + 0x00070007u;
    mn = pk16(lo16(mn) >> 4, hi16(mn) >> 4);
    mx = pk16(lo16(mx) >> 4, hi16(mx) >> 4);
    uint32_t vp = (uint32_t)hw | ((uint32_t)hh << 16);
    uint32_t vp2 = vp * 2;
    auto clampv = [&](uint32_t b) {
        int x = (int16_t)(lo16(b) + lo16(vp));
        int y = (int16_t)(hi16(b) + hi16(vp));
        x = std::min(std::max(x, 0), (int)lo16(vp2));
        y = std::min(std::max(y, 0), (int)hi16(vp2));
        return pk16(x, y);
    };
    mn = clampv(mn);
    mx = clampv(mx);
    bmin = mn & ~0x00030003u;
    bmax = (mx + 0x00030003u) & ~0x00030003u;
}

// TrianglePacket::Setup (Rasterizer.cpp:256-289). Returns the lane mask of
// front-facing triangles whose tile-aligned render box is not empty.
uint16_t setup_packet(TrianglePacket& t, const V4 v[3][16], int hw, int hh, int cullMode) {
    uint16_t mask = 0;
    float fsx = (float)(hw * 16), fsy = (float)(hh * 16);
    for (int l = 0; l < 16; l++) {
        float x[3], y[3], z[3], rw[3];
        for (int k = 0; k < 3; k++) {
            rw[k] = frcp(v[k][l].w);
            x[k] = v[k][l].x * rw[k];
            y[k] = v[k][l].y * rw[k];
            z[k] = v[k][l].z * rw[k];
        }
        // (x2-x0)*(y1-y0) - (x0-x1)*(y0-y2), contracted as in probe_setup
        float det = fmaf(x[2] - x[0], y[1] - y[0], (x[0] - x[1]) * (y[2] - y[0]));
        if (cullMode != 1) {
            bool flip = (cullMode == 2) ? true : (det < 0);
            det = flip ? -det : det;
        }
        bool keep = det > 0;
        int32_t X[3], Y[3];
        for (int k = 0; k < 3; k++) {
            X[k] = cvt_rne(x[k] * fsx);
            Y[k] = cvt_rne(y[k] * fsy);
        }
        t.Pos0[l] = ((uint32_t)X[0] & 0xFFFF) | ((uint32_t)Y[0] << 16);
        t.Pos1[l] = ((uint32_t)X[1] & 0xFFFF) | ((uint32_t)Y[1] << 16);
        t.Pos2[l] = ((uint32_t)X[2] & 0xFFFF) | ((uint32_t)Y[2] << 16);
        t.Z0[l] = z[0]; t.Z1[l] = z[1]; t.Z2[l] = z[2];
        t.W0[l] = rw[0]; t.W1[l] = rw[1]; t.W2[l] = rw[2];
        uint32_t bmin, bmax;
        render_bbox(t, l, hw, hh, bmin, bmax);
        if (lo16(bmin) >= lo16(bmax) || hi16(bmin) >= hi16(bmax)) keep = false;
        if (keep) mask |= (uint16_t)(1u << l);
    }
    return mask;
}

struct EdgeVars {  // Rasterizer.h:162-176
    int32_t Edge0[16], Edge1[16], Edge2[16];
    int32_t A01[16], A12[16], A20[16];
    int32_t B01[16], B12[16], B20[16];
    float Z0[16], Z10[16], Z20[16];
    float W0[16], W0S[16], W1S[16], W2S[16];
};

int32_t compute_edge(int32_t a, int32_t x, int32_t b, int32_t y) {  // Rasterizer.cpp:291-295
    int32_t w = wrap_add(wrap_mul(a, x), wrap_mul(b, y));
    w = wrap_add(w, (a > 0 || (a == 0 && b > 0)) ? 0 : -1);
    return w >> 4;
}

// TriangleEdgeVars::Setup (Rasterizer.cpp:296-329).
void setup_edges(EdgeVars& e, const TrianglePacket& t, int hw, int hh) {
    for (int l = 0; l < 16; l++) {
        int32_t x0 = lo16(t.Pos0[l]), y0 = (int32_t)t.Pos0[l] >> 16;
        int32_t x1 = lo16(t.Pos1[l]), y1 = (int32_t)t.Pos1[l] >> 16;
        int32_t x2 = lo16(t.Pos2[l]), y2 = (int32_t)t.Pos2[l] >> 16;
        int32_t A01 = y1 - y0, B01 = x0 - x1;
        int32_t A12 = y2 - y1, B12 = x1 - x2;
        int32_t A20 = y0 - y2, B20 = x2 - x0;
        int32_t det = wrap_sub(wrap_mul(B20, A01), wrap_mul(B01, A20));
        if (det < 0) {
            A01 = -A01; B01 = -B01; A12 = -A12; B12 = -B12; A20 = -A20; B20 = -B20;
            det = wrap_sub(0, det);
        }
        int32_t sx = -(hw << 4) + 8, sy = -(hh << 4) + 8;
        e.Edge0[l] = compute_edge(A12, sx - x1, B12, sy - y1);
        e.Edge1[l] = compute_edge(A20, sx - x2, B20, sy - y2);
        e.Edge2[l] = compute_edge(A01, sx - x0, B01, sy - y0);
        e.A01[l] = A01; e.A12[l] = A12; e.A20[l] = A20;
        e.B01[l] = B01; e.B12[l] = B12; e.B20[l] = B20;
        float rcpArea = fdiv(16.0f, (float)det);
        e.Z0[l] = t.Z0[l];
        e.Z10[l] = (t.Z1[l] - t.Z0[l]) * rcpArea;
        e.Z20[l] = (t.Z2[l] - t.Z0[l]) * rcpArea;
        e.W0[l] = t.W0[l];
        e.W0S[l] = t.W0[l] * rcpArea;
        e.W1S[l] = t.W1[l] * rcpArea;
        e.W2S[l] = t.W2[l] * rcpArea;
    }
}

inline uint32_t pixel_offset(uint32_t W, uint32_t x, uint32_t y) {  // Rasterizer.h:50-56
    return ((x & ~3u) << 2) + (y & ~3u) * W + (x & 3) + (y & 3) * 4;
}

struct DrawState {
    orc_framebuffer* fb;
    const orc_raster_opts* opts;
    orc_counters* ctr;
    int row0, row1;                // rows drawn: [row0, row1), multiples of 4 (the whole frame by default)
};

void sample_level16(const orc_texture* tex, const float* u, const float* v, int32_t layer, const int32_t* mip,
                    int32_t wrap, bool magLinear, bool minLinear, const bool* active, uint32_t* out);

// What FS_EncodeSurfaceId<true> (Shading.cpp:309-331) needs besides
// coverage and depth: the triangle's vertex UVs, its material's texture and
// alpha cutoff, and for clipped triangles the barycentric remap of the fan
// triangle onto the source triangle (TriangleClipData, Rasterizer.cpp:461-464).
struct FragmentCtx {
    bool alpha_test;
    float uv[3][2];
    const orc_texture* tex;
    uint32_t cutoff;
    bool clipped;
    float ru[3], rv[3];
    // FS_EncodeGBuffer: whether the meshlet has a material, the triangle's
    // vertex normal/tangent words and the handedness bit of its first vertex
    bool material;
    uint32_t nt[3];
    uint32_t handed;
    const float* object_to_world;
};

// Perspective-correct barycentrics of one fragment (Rasterizer.h:293-319):
// from the integer edge values u = E1, v = E2 and the triangle's 1/w terms.
void fragment_bary(float u, float v, float W0, float W0S, float W1S, float W2S, float& bu, float& bv) {
    float pw0 = fmaf(u + v, -W0S, W0);
    float w = fmaf(u, W1S, fmaf(v, W2S, pw0));
    float r = rcp14(w);
    r = r * fmaf(-w, r, 2.0f);
    // `u *= W1S * rcpW` is reassociated by the fast-math build to (u * W1S) * rcpW
    bu = (u * W1S) * r;
    bv = (v * W2S) * r;
}

// The alpha test of FS_EncodeSurfaceId<true> on one 4x4 tile: UVs of all 16
// pixels from their barycentrics (BaryLerp, Rasterizer.h:100-102),
// SampleImplicitLod (Texture.h:403-410: derivatives from pixel pairs of the
// tile, mip level less the 8 fraction bits, one filter per tile) and
// color >= cutoff << 24. Returns the passing pixels.
void alpha_uv_tile(const FragmentCtx& fc, float slu, float slv, const float* bu, const float* bv, float* U,
                   float* V, int32_t* mip, float (*bary)[16] = nullptr) {
    float su[16], sv[16];
    for (int l = 0; l < 16; l++) {
        float u = bu[l], v = bv[l];
        if (fc.clipped) {
            float cu = fmaf(u, fc.ru[1], fmaf(v, fc.ru[2], fc.ru[0]));
            float cv = fmaf(u, fc.rv[1], fmaf(v, fc.rv[2], fc.rv[0]));
            u = cu;
            v = cv;
        }
        float b0 = 1.0f - (u + v);   // `1 - u - v`, reassociated by the fast-math build
        if (bary) {
            bary[0][l] = b0;
            bary[1][l] = u;
            bary[2][l] = v;
        }
        U[l] = fmaf(fc.uv[0][0], b0, fmaf(fc.uv[1][0], u, fc.uv[2][0] * v));
        V[l] = fmaf(fc.uv[0][1], b0, fmaf(fc.uv[1][1], u, fc.uv[2][1] * v));
        su[l] = U[l] * slu;
        sv[l] = V[l] * slv;
    }
    for (int l = 0; l < 16; l++) {
        int a = l & ~1, b = l | 1;                   // dFdx: pixel pairs in a row
        int c = l & ~4, d = l | 4;                   // dFdy: row pairs
        float gx = su[b] - su[a], gy = sv[b] - sv[a];
        float gz = su[d] - su[c], gw = sv[d] - sv[c];
        float dx = fmaf(gx, gx, gy * gy), dy = fmaf(gz, gz, gw * gw);
        float mx = (dx > dy) ? dx : dy;
        mip[l] = (wrap_sub((int32_t)fbits(mx), 127 << 23) >> 24) - 8;
    }
}

uint16_t alpha_test_tile(const FragmentCtx& fc, const float* bu, const float* bv) {
    float U[16], V[16];
    int32_t mip[16];
    alpha_uv_tile(fc, fc.tex->scale_lerp_u, fc.tex->scale_lerp_v, bu, bv, U, V, mip);
    uint32_t color[16];
    const bool all[16] = { true, true, true, true, true, true, true, true,
                           true, true, true, true, true, true, true, true };
    sample_level16(fc.tex, U, V, 0, mip, 0, true, false, all, color);
    uint16_t pass = 0;
    for (int l = 0; l < 16; l++)
        if (color[l] >= (fc.cutoff << 24)) pass |= (uint16_t)(1u << l);
    return pass;
}

// FS_EncodeGBuffer (Shading.cpp:344-407) on one 4x4 tile after the depth
// test; defined with the surface helpers below.
void gbuffer_tile(DrawState& st, const EdgeVars& e, int l, int tx, int ty, uint16_t pass, const int32_t* E1,
                  const int32_t* E2, const float* depth, const FragmentCtx& fc);

// FS_Overdraw (Shading.cpp:333-342) on one 4x4 tile with at least one
// covered pixel: every pixel's counter gains 1 in its high 16 bits when
// covered, else 1 in its low 16 bits (saturating 16-bit adds), and every
// pixel's depth becomes max(plane depth, stored depth) -- uncovered pixels
// included, with the plane's value extrapolated to them. The max is
// vmaxps(depth, stored): the stored value unless the new one is greater.
void overdraw_tile(DrawState& st, int tx, int ty, uint16_t mask, const float* depth) {
    orc_framebuffer* fb = st.fb;
    uint32_t* counters = fb->data;
    float* depthBuf = (float*)(fb->data + fb->layer_stride);
    for (int p = 0; p < 16; p++) {
        uint32_t off = pixel_offset(fb->width, (uint32_t)(tx + (p & 3)), (uint32_t)(ty + (p >> 2)));
        uint32_t c = counters[off];
        uint32_t hi = c >> 16, lo = c & 0xFFFFu;
        if ((mask >> p) & 1) {
            hi = std::min(hi + 1u, 0xFFFFu);
            st.ctr->fragments_passed++;
        } else {
            lo = std::min(lo + 1u, 0xFFFFu);
        }
        counters[off] = (hi << 16) | lo;
        if (depth[p] > depthBuf[off]) depthBuf[off] = depth[p];
    }
}

// Rasterizer::DrawTriangle (Rasterizer.h:250-328) with FS_EncodeSurfaceId
// (Shading.cpp:309-331) inlined: per 4x4 tile the depth test, then, for
// alpha-tested materials, the alpha test. With the overdraw table the
// fragment shader is FS_Overdraw instead.
void draw_triangle(DrawState& st, const EdgeVars& e, int l, uint32_t bmin, uint32_t bmax, uint32_t surfaceId,
                   const FragmentCtx& fc) {
    orc_framebuffer* fb = st.fb;
    uint32_t W = fb->width;
    uint32_t* ids = fb->data;
    float* depthBuf = (float*)(fb->data + fb->layer_stride);
    int minX = lo16(bmin), minY = std::max((int)hi16(bmin), st.row0), maxX = lo16(bmax),
        maxY = std::min((int)hi16(bmax), st.row1);
    for (int ty = minY; ty < maxY; ty += 4) {
        for (int tx = minX; tx < maxX; tx += 4) {
            int32_t E1[16], E2[16];
            float depth[16];
            uint16_t mask = 0;
            for (int p = 0; p < 16; p++) {
                int x = tx + (p & 3), y = ty + (p >> 2);
                int32_t w0 = wrap_add(wrap_add(e.Edge0[l], wrap_mul(e.A12[l], x)), wrap_mul(e.B12[l], y));
                int32_t w1 = wrap_add(wrap_add(e.Edge1[l], wrap_mul(e.A20[l], x)), wrap_mul(e.B20[l], y));
                int32_t w2 = wrap_add(wrap_add(e.Edge2[l], wrap_mul(e.A01[l], x)), wrap_mul(e.B01[l], y));
                E1[p] = w1;
                E2[p] = w2;
                depth[p] = fmaf((float)w1, e.Z10[l], fmaf((float)w2, e.Z20[l], e.Z0[l]));
                if ((w0 | w1 | w2) < 0) continue;
                mask |= (uint16_t)(1u << p);
                st.ctr->fragments_covered++;
                if (st.opts && st.opts->coverage_count) st.opts->coverage_count[(size_t)y * W + x]++;
            }
            if (!mask) continue;
            if (st.opts && st.opts->pipeline == ORC_PIPELINE_OVERDRAW) {
                overdraw_tile(st, tx, ty, mask, depth);
                continue;
            }
            uint16_t pass = 0;
            for (int p = 0; p < 16; p++) {
                if (!((mask >> p) & 1)) continue;
                int x = tx + (p & 3), y = ty + (p >> 2);
                if (depth[p] > depthBuf[pixel_offset(W, (uint32_t)x, (uint32_t)y)]) pass |= (uint16_t)(1u << p);
            }
            if (!pass) continue;
            if (st.opts && st.opts->pipeline == ORC_PIPELINE_DEFERRED) {
                gbuffer_tile(st, e, l, tx, ty, pass, E1, E2, depth, fc);
                continue;
            }
            if (fc.alpha_test) {
                float bu[16], bv[16];
                for (int p = 0; p < 16; p++)
                    fragment_bary((float)E1[p], (float)E2[p], e.W0[l], e.W0S[l], e.W1S[l], e.W2S[l], bu[p], bv[p]);
                pass &= alpha_test_tile(fc, bu, bv);
            }
            for (int p = 0; p < 16; p++) {
                if (!((pass >> p) & 1)) continue;
                int x = tx + (p & 3), y = ty + (p >> 2);
                uint32_t off = pixel_offset(W, (uint32_t)x, (uint32_t)y);
                depthBuf[off] = depth[p];
                ids[off] = surfaceId;
                st.ctr->fragments_passed++;
            }
        }
    }
}

// ---------------------------------------------------------------- clipper

// Clipper (Rasterizer.cpp:56-82, 398-491).
struct Clipper {
    float bx, by;
    float Vertices[6][16 * 3 + 16];
    V4 Out[3][8][16];
    uint8_t PrimId[8][16];
    float ClipU[8][16][3], ClipV[8][16][3];   // TriangleClipData::ClippedU/V
    uint16_t Masks[8];
    TrianglePacket Tris[8];
    uint32_t OutIdx = 0;

    float dist(uint8_t i, uint32_t plane, float scale) const {
        float a = Vertices[plane / 2][i];
        if (plane % 2) a = -a;
        return fmaf(Vertices[3][i], scale, a);  // a + w*scale, contracted
    }

    void flush(int hw, int hh, int cullMode, bool isLast) {
        uint32_t pi = OutIdx / 16;
        V4 v[3][16];
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 16; l++) v[k][l] = Out[k][pi][l];
        uint16_t m = setup_packet(Tris[pi], v, hw, hh, cullMode);
        m &= (uint16_t)((1u << (OutIdx % 16 + (isLast ? 0 : 1))) - 1);
        Masks[pi] = m;
    }

    void clip(const V4 vin[3][16], const ClipCodes& cc, int hw, int hh, int cullMode) {
        for (int k = 0; k < 3; k++) {
            for (int l = 0; l < 16; l++) {
                int j = k * 16 + l;
                Vertices[0][j] = vin[k][l].x;
                Vertices[1][j] = vin[k][l].y;
                Vertices[2][j] = vin[k][l].z;
                Vertices[3][j] = vin[k][l].w;
                Vertices[4][j] = k == 1 ? 1.0f : 0.0f;
                Vertices[5][j] = k == 2 ? 1.0f : 0.0f;
            }
        }
        for (int prim = 0; prim < 16; prim++) {
            if (!((cc.NonTrivialMask >> prim) & 1)) continue;
            uint8_t idx[32] = { (uint8_t)prim, (uint8_t)(prim + 16), (uint8_t)(prim + 32) };
            uint32_t vertCount = 3, nextIdx = 48;
            for (uint32_t plane = 0; plane < 8; plane++) {
                if (!((cc.OutCodes[prim] >> plane) & 1)) continue;
                uint8_t out[32];
                uint32_t outCount = 0;
                float scale = plane < 4 ? (plane / 2 == 0 ? bx : by) : 1.0f;
                for (uint32_t vi = 0; vi < vertCount; vi++) {
                    uint8_t ia = idx[vi], ib = idx[(vi + 1) % vertCount];
                    float da = dist(ia, plane, scale), db = dist(ib, plane, scale);
                    if (da >= 0) out[outCount++] = ia;
                    // the vertex store has room for 16 generated vertices
                    // per triangle (Rasterizer.cpp:64, 441)
                    if ((da >= 0) != (db >= 0) && nextIdx < 64) {
                        float t = da / (da - db);
                        for (int k = 0; k < 6; k++)
                            Vertices[k][nextIdx] = fmaf(Vertices[k][ib], t, fmaf(-t, Vertices[k][ia], Vertices[k][ia]));
                        out[outCount++] = (uint8_t)nextIdx++;
                    }
                }
                if (vertCount < 3) break;
                vertCount = outCount;
                memcpy(idx, out, sizeof(idx));
            }
            if (vertCount < 2) continue;
            // at most 8 fan triangles per source triangle (a clipped
            // triangle has at most 9 vertices, i.e. 7 fan triangles)
            for (uint32_t vi = 0; vi + 2 < vertCount && vi < 8; vi++) {
                uint32_t i0 = idx[0], i1 = idx[vi + 1], i2 = idx[vi + 2];
                uint32_t lane = OutIdx % 16, pi = OutIdx / 16;
                uint32_t ii[3] = { i0, i1, i2 };
                for (int k = 0; k < 3; k++)
                    Out[k][pi][lane] = { Vertices[0][ii[k]], Vertices[1][ii[k]], Vertices[2][ii[k]], Vertices[3][ii[k]] };
                PrimId[pi][lane] = (uint8_t)prim;
                const float u0 = Vertices[4][i0], v0 = Vertices[5][i0];
                ClipU[pi][lane][0] = u0;
                ClipU[pi][lane][1] = Vertices[4][i1] - u0;
                ClipU[pi][lane][2] = Vertices[4][i2] - u0;
                ClipV[pi][lane][0] = v0;
                ClipV[pi][lane][1] = Vertices[5][i1] - v0;
                ClipV[pi][lane][2] = Vertices[5][i2] - v0;
                if (lane == 15) flush(hw, hh, cullMode, false);
                if (++OutIdx >= 128) return;
            }
        }
        if (OutIdx % 16 != 0) flush(hw, hh, cullMode, true);
    }
};

// ---------------------------------------------------------------- textures

struct Tex { const orc_texture* t; };

inline uint32_t texel_offset(int32_t x, int32_t y, int32_t stride) {  // TiledY8, Texture.h:482-491
    return (uint32_t)((y & 7) | (x << 3) | ((y & ~7) << stride));
}

inline uint32_t lerp16(uint32_t a, uint32_t b, uint32_t t) {  // SIMD.h:450 per 16-bit lane
    uint32_t r = 0;
    for (int h = 0; h < 2; h++) {
        int16_t av = (int16_t)(a >> (16 * h)), bv = (int16_t)(b >> (16 * h)), tv = (int16_t)(t >> (16 * h));
        int16_t d = (int16_t)(bv - av);
        int32_t prod = (int32_t)d * (int32_t)tv;
        int16_t mh = (int16_t)(((prod >> 14) + 1) >> 1);  // mulhrs
        r |= (uint32_t)(uint16_t)(int16_t)(av + mh) << (16 * h);
    }
    return r;
}

// Texture2D::SampleLevel for RGBA8u (Texture.h:412-459, 506-556). The filter
// is chosen once for the 16 lanes of a 4x4 tile (Texture.h:432).
void sample_level16(const orc_texture* tex, const float* u, const float* v, int32_t layer,
                    const int32_t* mipIn, int wrap, bool magLinear, bool minLinear,
                    const bool* active, uint32_t* out) {
    int32_t ix[16], iy[16];
    bool anyPos = false;
    for (int l = 0; l < 16; l++) {
        float su = u[l] * tex->scale_lerp_u, sv = v[l] * tex->scale_lerp_v;
        ix[l] = cvt_rne(su);
        iy[l] = cvt_rne(sv);
        if (wrap == 0) {  // Repeat
            ix[l] &= tex->mask_lerp_u;
            iy[l] &= tex->mask_lerp_v;
        } else if (wrap == 1) {  // ClampToEdge
            ix[l] = std::min(std::max(ix[l], 0), tex->mask_lerp_u);
            iy[l] = std::min(std::max(iy[l], 0), tex->mask_lerp_v);
        } else {  // MirroredRepeat
            ix[l] = (ix[l] & tex->mask_lerp_u) ^ ((ix[l] & (tex->mask_lerp_u + 1)) != 0 ? tex->mask_lerp_u : 0);
            iy[l] = (iy[l] & tex->mask_lerp_v) ^ ((iy[l] & (tex->mask_lerp_v + 1)) != 0 ? tex->mask_lerp_v : 0);
        }
        if ((!active || active[l]) && mipIn[l] > 0) anyPos = true;
    }
    bool linear = anyPos ? minLinear : magLinear;
    int maxLevel = (int)tex->mip_levels - 1;
    for (int l = 0; l < 16; l++) {
        int m = std::min(std::max(mipIn[l], 0), maxLevel);
        int32_t offset = layer * (int32_t)tex->layer_stride + (int32_t)tex->mip_offsets[m];
        int32_t stride = (int32_t)tex->row_shift - m;
        int32_t x = ix[l] >> m, y = iy[l] >> m;
        const uint32_t* D = tex->data;
        if (!linear) {
            out[l] = D[offset + (int32_t)texel_offset(x >> 8, y >> 8, stride)];
            continue;
        }
        int32_t xf = std::max(x - 127, 0), yf = std::max(y - 127, 0);
        int32_t tx = xf >> 8, ty = yf >> 8;
        bool inX = ((tx + 1) << m) < (int32_t)tex->width;
        bool inY = ((ty + 1) << m) < (int32_t)tex->height;
        int32_t i00 = offset + (int32_t)texel_offset(tx, ty, stride);
        int32_t i01 = offset + (int32_t)texel_offset(tx, ty + (inY ? 1 : 0), stride);
        uint32_t d00 = D[i00], d10 = D[i00 + 8], d01 = D[i01], d11 = D[i01 + 8];
        uint32_t fx = (uint32_t)(xf & 255) << 7, fy = (uint32_t)(yf & 255) << 7;
        fx = (fx << 16) | fx;
        fy = (fy << 16) | fy;
        if (!inX) fx = 0;
        uint32_t rb1 = lerp16(d00 & 0x00FF00FF, d10 & 0x00FF00FF, fx);
        uint32_t ga1 = lerp16((d00 >> 8) & 0x00FF00FF, (d10 >> 8) & 0x00FF00FF, fx);
        uint32_t rb2 = lerp16(d01 & 0x00FF00FF, d11 & 0x00FF00FF, fx);
        uint32_t ga2 = lerp16((d01 >> 8) & 0x00FF00FF, (d11 >> 8) & 0x00FF00FF, fx);
        uint32_t rb = lerp16(rb1, rb2, fy), ga = lerp16(ga1, ga2, fy);
        out[l] = rb | (ga << 8);
    }
}

// ---------------------------------------------------------------- resolve

// IntersectTriangle (Shading.cpp:417-464) in the operation order clang's
// fast-math build uses (probe_intersect).
void intersect_triangle(const float c[3][4], float su, float sv, float sw, float sh,
                        float bary[3], float ddx[3], float ddy[3]) {
    float iw0 = frcp(c[0][3]), iw1 = frcp(c[1][3]), iw2 = frcp(c[2][3]);
    float p0x = c[0][0] * iw0, p0y = c[0][1] * iw0;
    float p1x = c[1][0] * iw1, p1y = c[1][1] * iw1;
    float p2x = c[2][0] * iw2, p2y = c[2][1] * iw2;
    float m0x = p2x - p1x, m1x = p0x - p1x, m1y = p0y - p1y;
    float d12y = p1y - p2y;
    float det = fmaf(m0x, m1y, d12y * m1x);
    float invDet = frcp(det);
    float a0 = invDet * iw0, a1 = invDet * iw1, a2 = invDet * iw2;
    float dx0 = a0 * d12y, dx1 = a1 * (p2y - p0y), dx2 = a2 * m1y;
    float dy0 = a0 * m0x, dy1 = a1 * (p0x - p2x), dy2 = a2 * (p1x - p0x);
    float dxs = (dx0 + dx1) + dx2, dys = (dy0 + dy1) + dy2;
    float r0x = su - p0x, r0y = sv - p0y;
    float iiw = fmaf(dys, r0y, fmaf(dxs, r0x, iw0));
    float iW = frcp(iiw);
    float by = iW * fmaf(dx1, r0x, dy1 * r0y);
    float bz = iW * fmaf(dx2, r0x, dy2 * r0y);
    float bx = 1.0f - (by + bz);
    float kx = 2.0f / sw, ky = -2.0f / sh;
    float wdx = frcp(fmaf(kx, dxs, iiw)), wdy = frcp(fmaf(ky, dys, iiw));
    float b[3] = { bx, by, bz }, dxk[3] = { dx0, dx1, dx2 }, dyk[3] = { dy0, dy1, dy2 };
    for (int k = 0; k < 3; k++) {
        float bi = b[k] * iiw;
        ddx[k] = fmaf(wdx, fmaf(kx, dxk[k], bi), -b[k]);
        ddy[k] = fmaf(wdy, fmaf(ky, dyk[k], bi), -b[k]);
        bary[k] = b[k];
    }
}

// IntersectTriangle's operation order with correctly rounded divisions (the
// differentiable resolve's barycentrics).
void intersect_triangle_exact(const float c[3][4], float su, float sv, float sw, float sh,
                              float bary[3], float ddx[3], float ddy[3]) {
    float iw0 = 1.0f / c[0][3], iw1 = 1.0f / c[1][3], iw2 = 1.0f / c[2][3];
    float p0x = c[0][0] * iw0, p0y = c[0][1] * iw0;
    float p1x = c[1][0] * iw1, p1y = c[1][1] * iw1;
    float p2x = c[2][0] * iw2, p2y = c[2][1] * iw2;
    float m0x = p2x - p1x, m1x = p0x - p1x, m1y = p0y - p1y;
    float d12y = p1y - p2y;
    float det = fmaf(m0x, m1y, d12y * m1x);
    float invDet = 1.0f / det;
    float a0 = invDet * iw0, a1 = invDet * iw1, a2 = invDet * iw2;
    float dx0 = a0 * d12y, dx1 = a1 * (p2y - p0y), dx2 = a2 * m1y;
    float dy0 = a0 * m0x, dy1 = a1 * (p0x - p2x), dy2 = a2 * (p1x - p0x);
    float dxs = (dx0 + dx1) + dx2, dys = (dy0 + dy1) + dy2;
    float r0x = su - p0x, r0y = sv - p0y;
    float iiw = fmaf(dys, r0y, fmaf(dxs, r0x, iw0));
    float iW = 1.0f / iiw;
    float by = iW * fmaf(dx1, r0x, dy1 * r0y);
    float bz = iW * fmaf(dx2, r0x, dy2 * r0y);
    float bx = 1.0f - (by + bz);
    float kx = 2.0f / sw, ky = -2.0f / sh;
    float wdx = 1.0f / fmaf(kx, dxs, iiw), wdy = 1.0f / fmaf(ky, dys, iiw);
    float b[3] = { bx, by, bz }, dxk[3] = { dx0, dx1, dx2 }, dyk[3] = { dy0, dy1, dy2 };
    for (int k = 0; k < 3; k++) {
        float bi = b[k] * iiw;
        ddx[k] = fmaf(wdx, fmaf(kx, dxk[k], bi), -b[k]);
        ddy[k] = fmaf(wdy, fmaf(ky, dyk[k], bi), -b[k]);
        bary[k] = b[k];
    }
}

struct SurfaceLane {
    bool active;
    float u, v;
    int32_t mip;       // per material texture scale
    uint32_t material;
    float bary[3];
    float nrm[3];
    float tan[3];      // world-space tangent (lit resolve)
    uint32_t handed;   // 1 << 31 when vertex 0's tangent handedness bit is set
    uint32_t albedo;
    uint32_t nmr;      // layer 1 texel (normal xy, metallic, roughness), 0 without
};

// ResolveSurface (Shading.cpp:471-579) for one 4x4 tile: vertex fetch,
// re-transform, analytic barycentrics, UV gradients and albedo sampling.
// The tile-wide filter choice is taken over the tile's covered lanes; the
// reference also lets uncovered (sky) lanes vote with register contents left
// over from the waterfall loop, which this restatement leaves out.
// RGBA8u::Unpack (Texture.h:28-36).
void unpack_rgba8(uint32_t p, float c[4]) {
    const float scale = 1.0f / 255;
    for (int k = 0; k < 4; k++) c[k] = (float)((p >> (8 * k)) & 255) * scale;
}

inline float dot3(const float a[3], const float b[3]) {  // SIMD.h:429
    return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2]));
}

// normalize via VRSQRT14 (SIMD.h:439)
void normalize3(float v[3]) {
    float r = rsqrt14(dot3(v, v));
    for (int k = 0; k < 3; k++) v[k] = v[k] * r;
}

// y < 0 ? -x : x by sign bit (SIMD.h:341-344)
inline float mulsign(float x, float y) { return bitsf(fbits(x) ^ (fbits(y) & 0x80000000u)); }

// UnmapOctahedron (Texture.h:289-296).
void unmap_octahedron(float u, float v, float n[3]) {
    u = u * 2.0f - 1.0f;
    v = v * 2.0f - 1.0f;
    n[0] = u;
    n[1] = v;
    n[2] = (1.0f - fabsf(u)) - fabsf(v);
    float t = fmaxf(-n[2], 0.0f);
    n[0] = n[0] - mulsign(t, n[0]);
    n[1] = n[1] - mulsign(t, n[1]);
    normalize3(n);
}

// UnpackNormalTangent (Shading.cpp:232-236): normal from bytes 0-1, tangent
// from bytes 2-3.
void unpack_normal(uint32_t packed, float n[3]) {
    float c[4];
    unpack_rgba8(packed, c);
    unmap_octahedron(c[0], c[1], n);
}
void unpack_tangent(uint32_t packed, float t[3]) {
    float c[4];
    unpack_rgba8(packed, c);
    unmap_octahedron(c[2], c[3], t);
}

// normalize(mul(ObjectToWorldMat, BaryLerp(bary, a0, a1, a2)))
// (Shading.cpp:546-554, BaryLerp Rasterizer.h:101-104)
void lerp_to_world(const float a[3][3], const float bary[3], const float* M, float out[3]) {
    float ai[3];
    for (int j = 0; j < 3; j++) ai[j] = fmaf(a[0][j], bary[0], fmaf(a[1][j], bary[1], a[2][j] * bary[2]));
    for (int j = 0; j < 3; j++) out[j] = fmaf(ai[0], M[0 * 3 + j], fmaf(ai[1], M[1 * 3 + j], ai[2] * M[2 * 3 + j]));
    normalize3(out);
}

// Normal mapping of ResolveSurface (Shading.cpp:556-572) for one lane:
// bitangent = cross(N, T) with the handedness sign (SIMD.h:431-437),
// tangent-space normal from the texel's two 8-bit channels with z =
// approx_sqrt(1 - x^2 - y^2) (SIMD.h:296), result normalize(xT + yB + zN).
// The fast-math build contracts `c * (1/127.5) - 1` and the sums into FMAs
// and evaluates 1 - (x^2 + y^2) as 1 + fma(-x, x, -y^2).
void normal_map(const float n[3], const float t[3], uint32_t handed, uint32_t nmr, float out[3]) {
    float b[3] = { fmaf(n[1], t[2], -(n[2] * t[1])), fmaf(n[2], t[0], -(n[0] * t[2])),
                   fmaf(n[0], t[1], -(n[1] * t[0])) };
    for (int k = 0; k < 3; k++) b[k] = bitsf(fbits(b[k]) ^ handed);
    const float sx = fmaf((float)(nmr & 255), 1.0f / 127.5f, -1.0f);
    const float sy = fmaf((float)((nmr >> 8) & 255), 1.0f / 127.5f, -1.0f);
    const float s = 1.0f + fmaf(-sx, sx, -(sy * sy));
    const float sz = rsqrt14(s) * s;
    float r[3];
    for (int k = 0; k < 3; k++) r[k] = fmaf(sz, n[k], fmaf(sy, b[k], sx * t[k]));
    const float inv = rsqrt14(fmaf(r[0], r[0], fmaf(r[1], r[1], r[2] * r[2])));
    for (int k = 0; k < 3; k++) out[k] = r[k] * inv;
}

// cvttps2udq: truncation; NaN and results outside [0, 2^32) give 0xFFFFFFFF.
uint32_t cvt_trunc_u32(float x) {
    if (!(x > -1.0f && x < 4294967296.0f)) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// The layer-2 word of FS_EncodeGBuffer (Shading.cpp:397-401): MapOctahedron
// (Texture.h:282-288) of the remapped normal, each coordinate as
// conv<uint32_t>(oct * 1023 + 0.5) in 10 bits, then the NMR texel's
// metallic and roughness bytes cut to 6 bits. The fast-math build folds
// (c * 0.5 + 0.5) * 1023 + 0.5 into fma(c, 511.5, 512), sums the absolute
// values left to right, takes 1/sum from VRCP14PS alone and evaluates
// max(-z * w, 0) as vmaxps (zero unless positive).
uint32_t gbuffer_pack(const float n[3], uint32_t nmr) {
    const float w = rcp14((fabsf(n[0]) + fabsf(n[1])) + fabsf(n[2]));
    const float mz = -n[2] * w;
    const float t = mz > 0.0f ? mz : 0.0f;
    const float u = fmaf(n[0], w, mulsign(t, n[0]));
    const float v = fmaf(n[1], w, mulsign(t, n[1]));
    uint32_t r = cvt_trunc_u32(fmaf(u, 511.5f, 512.0f));
    r |= cvt_trunc_u32(fmaf(v, 511.5f, 512.0f)) << 10;
    r |= ((nmr >> 18) & 0x3Fu) << 20;
    r |= ((nmr >> 26) & 0x3Fu) << 26;
    return r;
}

void gbuffer_tile(DrawState& st, const EdgeVars& e, int l, int tx, int ty, uint16_t pass, const int32_t* E1,
                  const int32_t* E2, const float* depth, const FragmentCtx& fc) {
    orc_framebuffer* fb = st.fb;
    const uint32_t W = fb->width;
    float* depthBuf = (float*)(fb->data + fb->layer_stride);
    uint32_t* color = fb->data;
    uint32_t* ch2Buf = fb->data + 2 * (size_t)fb->layer_stride;
    uint32_t base[16] = {}, ch2[16] = {};
    if (fc.material && fc.tex) {
        // BaryLerp of the vertex UVs and SampleImplicitLod, as the alpha test
        float bu[16], bv[16], U[16], V[16], bary[3][16];
        int32_t mip[16];
        for (int p = 0; p < 16; p++)
            fragment_bary((float)E1[p], (float)E2[p], e.W0[l], e.W0S[l], e.W1S[l], e.W2S[l], bu[p], bv[p]);
        alpha_uv_tile(fc, fc.tex->scale_lerp_u, fc.tex->scale_lerp_v, bu, bv, U, V, mip, bary);
        const bool all[16] = { true, true, true, true, true, true, true, true,
                               true, true, true, true, true, true, true, true };
        sample_level16(fc.tex, U, V, 0, mip, 0, true, false, all, base);
        for (int p = 0; p < 16; p++)
            if (base[p] < (fc.cutoff << 24)) pass &= (uint16_t)~(1u << p);
        if (!pass) return;
        if (fc.tex->num_layers >= 2) {
            uint32_t nmr[16];
            sample_level16(fc.tex, U, V, 1, mip, 0, true, false, all, nmr);
            float nv[3][3], tv[3][3];
            for (int k = 0; k < 3; k++) {
                unpack_normal(fc.nt[k], nv[k]);
                unpack_tangent(fc.nt[k], tv[k]);
            }
            for (int p = 0; p < 16; p++) {
                if (!((pass >> p) & 1)) continue;
                const float b[3] = { bary[0][p], bary[1][p], bary[2][p] };
                float N[3], T[3], R[3];
                lerp_to_world(nv, b, fc.object_to_world, N);
                lerp_to_world(tv, b, fc.object_to_world, T);
                normal_map(N, T, fc.handed, nmr[p], R);
                ch2[p] = gbuffer_pack(R, nmr[p]);
            }
        }
    }
    for (int p = 0; p < 16; p++) {
        if (!((pass >> p) & 1)) continue;
        const uint32_t off = pixel_offset(W, (uint32_t)(tx + (p & 3)), (uint32_t)(ty + (p >> 2)));
        depthBuf[off] = depth[p];
        // a meshlet without a material writes depth only (Shading.cpp:352-355)
        if (fc.material && fc.tex) {
            color[off] = base[p];
            ch2Buf[off] = ch2[p];
        }
        st.ctr->fragments_passed++;
    }
}

void resolve_surface_tile(const orc_shading_ctx& ctx, const uint32_t* ids, const bool* active,
                          const float* suv_x, const float* suv_y, float sw, float sh,
                          SurfaceLane* lanes, const float* object_to_world = nullptr) {
    const orc_meshlets& ms = *ctx.meshlets;
    float gradU[16][2], gradV[16][2];  // (dx, dy) of u and v
    for (int l = 0; l < 16; l++) {
        SurfaceLane& L = lanes[l];
        L.active = active[l];
        L.albedo = 0;
        L.nmr = 0;
        L.handed = 0;
        L.material = UINT32_MAX;
        L.u = L.v = 0;
        gradU[l][0] = gradU[l][1] = gradV[l][0] = gradV[l][1] = 0;
        if (!active[l]) continue;
        uint32_t sid = ids[l];
        uint32_t mi = sid / 128, tri = sid % 128;
        float c[3][4];
        float uv[3][2];
        for (int k = 0; k < 3; k++) {
            uint32_t vi = ms.indices[(size_t)mi * 384 + k * 128 + tri];
            const float* P = ms.positions + (size_t)mi * 192;
            transform(ctx.object_to_clip, P[vi], P[64 + vi], P[128 + vi], c[k]);
            uint32_t tc = ms.texcoords[(size_t)mi * 64 + vi];
            uv[k][0] = half_to_float((uint16_t)(tc & 0xFFFF));
            uv[k][1] = half_to_float((uint16_t)(tc >> 16));
        }
        L.material = ms.material_id[mi];
        float bary[3], ddx[3], ddy[3];
        intersect_triangle(c, suv_x[l], suv_y[l], sw, sh, bary, ddx, ddy);
        for (int k = 0; k < 3; k++) L.bary[k] = bary[k];
        if (object_to_world) {
            // normalWS = normalize(mul(ObjectToWorldMat, BaryLerp(bary, n0, n1, n2)))
            // (Shading.cpp:546-552, BaryLerp Rasterizer.h:101-104)
            float nv[3][3], tv[3][3];
            for (int k = 0; k < 3; k++) {
                uint32_t vi = ms.indices[(size_t)mi * 384 + k * 128 + tri];
                unpack_normal(ms.normal_tangents[(size_t)mi * 64 + vi], nv[k]);
                unpack_tangent(ms.normal_tangents[(size_t)mi * 64 + vi], tv[k]);
            }
            lerp_to_world(nv, bary, object_to_world, L.nrm);
            lerp_to_world(tv, bary, object_to_world, L.tan);
            // handedness of the first vertex (Shading.cpp:501-503)
            uint32_t v0 = ms.indices[(size_t)mi * 384 + tri] & 63u;
            if ((ms.tangent_handedness[mi] >> v0) & 1u) L.handed = 1u << 31;
        }
        float t10x = uv[1][0] - uv[0][0], t10y = uv[1][1] - uv[0][1];
        float t20x = uv[2][0] - uv[0][0], t20y = uv[2][1] - uv[0][1];
        L.u = fmaf(bary[2], t20x, fmaf(bary[1], t10x, uv[0][0]));
        L.v = fmaf(bary[2], t20y, fmaf(bary[1], t10y, uv[0][1]));
        gradU[l][0] = fmaf(t20x, ddx[2], t10x * ddx[1]);
        gradV[l][0] = fmaf(t20y, ddx[2], t10y * ddx[1]);
        gradU[l][1] = fmaf(t20x, ddy[2], t10x * ddy[1]);
        gradV[l][1] = fmaf(t20y, ddy[2], t10y * ddy[1]);
    }
    // Material waterfall (Shading.cpp:532-545): for each distinct material in
    // the tile, the mip level of every lane is taken with that material's
    // texture scale, and the filter choice is made over all covered lanes.
    bool done[16] = {};
    for (int l0 = 0; l0 < 16; l0++) {
        if (!active[l0] || done[l0] || lanes[l0].material == UINT32_MAX) continue;
        uint32_t mat = lanes[l0].material;
        const orc_material& mt = ctx.materials[mat];
        if (mt.texture < 0) {
            for (int l = 0; l < 16; l++) if (active[l] && lanes[l].material == mat) done[l] = true;
            continue;
        }
        const orc_texture* tex = &ctx.textures[mt.texture];
        float ssu = tex->scale_u * tex->scale_u, ssv = tex->scale_v * tex->scale_v;
        int32_t mip[16];
        float U[16], V[16];
        for (int l = 0; l < 16; l++) {
            float dx = fmaf(gradU[l][0], gradU[l][0], gradV[l][0] * gradV[l][0]) * ssu;
            float dy = fmaf(gradU[l][1], gradU[l][1], gradV[l][1] * gradV[l][1]) * ssv;
            float mx = (dx > dy) ? dx : dy;  // vmaxps: second operand on NaN/equal
            mip[l] = wrap_sub((int32_t)fbits(mx), 127 << 23) >> 24;
            U[l] = lanes[l].u;
            V[l] = lanes[l].v;
        }
        uint32_t res[16], nmr[16] = {};
        // SurfaceSampler: Repeat, Mag Linear, Min Nearest (Shading.cpp:6-10)
        sample_level16(tex, U, V, 0, mip, 0, true, false, active, res);
        // the normal/metallic/roughness layer (Shading.cpp:541-543)
        if (object_to_world && tex->num_layers >= 2) sample_level16(tex, U, V, 1, mip, 0, true, false, active, nmr);
        for (int l = 0; l < 16; l++) {
            if (active[l] && lanes[l].material == mat) {
                lanes[l].albedo = res[l];
                lanes[l].nmr = nmr[l];
                done[l] = true;
            }
        }
    }
    // normal mapping when a lane of the tile has a normal texel
    // (Shading.cpp:556-572)
    bool anyNormal = false;
    for (int l = 0; l < 16; l++) anyNormal |= (lanes[l].nmr & 0xFFFFu) != 0;
    if (object_to_world && anyNormal) {
        for (int l = 0; l < 16; l++) {
            if (!active[l]) continue;
            float n[3] = { lanes[l].nrm[0], lanes[l].nrm[1], lanes[l].nrm[2] };
            normal_map(n, lanes[l].tan, lanes[l].handed, lanes[l].nmr, lanes[l].nrm);
        }
    }
}

// RGBA8u::UnpackSrgb (Texture.h:37-54): squares the byte value with a
// 16-bit high multiply; alpha keeps the linear value.
void unpack_srgb(uint32_t packed, float c[4]) {
    uint32_t rb1 = ((packed << 8) & 0xFF00FF00u) + 0x00FF00FFu;
    uint32_t ag1 = (packed & 0xFF00FF00u) + 0x00FF00FFu;
    auto mulhi16 = [](uint32_t a) {
        uint32_t lo = a & 0xFFFFu, hi = a >> 16;
        return ((lo * lo) >> 16) | (((hi * hi) >> 16) << 16);
    };
    uint32_t rb2 = mulhi16(rb1), ag2 = mulhi16(ag1);
    const float scale = 1.0f / 65535;
    c[0] = (float)(rb2 & 65535) * scale;
    c[1] = (float)(ag2 & 65535) * scale;
    c[2] = (float)(rb2 >> 16) * scale;
    c[3] = (float)(ag1 >> 16) * scale;
}

inline float lerpf(float a, float b, float t) { return fmaf(t, b, fmaf(-t, a, a)); }  // SIMD.h:446

// Float min/max/clamp are maxnum/minnum (SIMD.h:319-331): a NaN operand
// yields the other one, hence fmaxf/fminf rather than std::max/min.

float light_attenuation(const orc_light& L, const float pos[3]) {  // Shading.cpp:581-600
    if (L.type == 0) return 1.0f;
    float d[3] = { L.position[0] - pos[0], L.position[1] - pos[1], L.position[2] - pos[2] };
    float d2 = dot3(d, d);
    float factor = d2 * L.inv_radius_sq;
    float smooth = fmaxf(1.0f - factor * factor, 0.0f);
    float att = (smooth * smooth) * rcp14(fmaxf(d2, 1e-4f));
    if (L.type == 2) {
        float nd[3] = { d[0], d[1], d[2] };
        normalize3(nd);
        float negDir[3] = { -L.direction[0], -L.direction[1], -L.direction[2] };
        float cd = dot3(negDir, nd);
        float sa = fminf(fmaxf(cd * L.spot_scale + L.spot_offset, 0.0f), 1.0f);
        att *= sa * sa;
    }
    return att;
}

// EvalLighting for one 4x4 tile (Shading.cpp:602-645): Lambert diffuse,
// GGX/Smith/Schlick specular, 5% ambient. The per-light skips are taken over
// the tile's covered lanes, as the reference takes them over its 16 lanes.
void lit_tile_at(const SurfaceLane* lanes, const bool* active, const float (*wpos)[3], const orc_resolve_params* p,
                 float exposure, float color[16][3]) {
    float base[16][3], f0[16][3], diffuse[16][3], N[16][3], V[16][3], NoV[16], aR[16], metal[16];
    for (int l = 0; l < 16; l++) {
        if (!active[l]) continue;
        float srgb[4];
        unpack_srgb(lanes[l].albedo, srgb);
        // SurfaceData::Metallic / Roughness (Shading.cpp:574-577), 0 without an NMR layer
        metal[l] = (float)((lanes[l].nmr >> 16) & 255) * (1.0f / 255);
        float rough = (float)((lanes[l].nmr >> 24) & 255) * (1.0f / 255);
        aR[l] = fmaxf(rough * rough, 1e-4f);
        const float refl = 0.16f * 0.5f * 0.5f;
        for (int k = 0; k < 3; k++) {
            base[l][k] = srgb[k];
            f0[l][k] = lerpf(refl, base[l][k], metal[l]);
            diffuse[l][k] = base[l][k] * (1.0f - metal[l]);
            N[l][k] = lanes[l].nrm[k];
            V[l][k] = p->view_pos[k] - wpos[l][k];
        }
        normalize3(V[l]);
        NoV[l] = fabsf(dot3(N[l], V[l])) + 1e-5f;
        for (int k = 0; k < 3; k++) color[l][k] = 0.0f;
    }
    for (uint32_t li = 0; li < p->num_lights; li++) {
        const orc_light& L = p->lights[li];
        float Ld[16][3], NoL[16], att[16];
        bool allLow = true;
        for (int l = 0; l < 16; l++) {
            if (!active[l]) continue;
            if (L.type == 0) {
                for (int k = 0; k < 3; k++) Ld[l][k] = -L.direction[k];
            } else {
                for (int k = 0; k < 3; k++) Ld[l][k] = L.position[k] - wpos[l][k];
                normalize3(Ld[l]);
            }
            NoL[l] = dot3(N[l], Ld[l]);
            allLow &= NoL[l] < 1e-4f;
        }
        if (allLow) continue;
        bool allDark = true;
        for (int l = 0; l < 16; l++) {
            if (!active[l]) continue;
            att[l] = light_attenuation(L, wpos[l]) * L.intensity * exposure;
            allDark &= NoL[l] * att[l] < 1e-4f;
        }
        if (allDark) continue;
        for (int l = 0; l < 16; l++) {
            if (!active[l]) continue;
            float Hv[3] = { V[l][0] + Ld[l][0], V[l][1] + Ld[l][1], V[l][2] + Ld[l][2] };
            normalize3(Hv);
            float NoH = fminf(fmaxf(dot3(N[l], Hv), 0.0f), 1.0f);
            float LoH = fminf(fmaxf(dot3(Ld[l], Hv), 0.0f), 1.0f);
            // D_GGX, V_SmithGGXCorrelatedFast, F_Schlick (Shading.cpp:19-32)
            float a = NoH * aR[l];
            float k = aR[l] * rcp14((1.0f - NoH * NoH) + a * a);
            float D = k * k * 0.3183098861837907f;
            float va = 2.0f * NoL[l] * NoV[l];
            float vb = NoL[l] + NoV[l];
            float Vis = fdiv(0.5f, lerpf(va, vb, aR[l]));
            float u1 = 1.0f - LoH;
            float f = (u1 * u1) * (u1 * u1) * u1;
            float w = fmaxf(NoL[l] * att[l], 0.0f);
            for (int c = 0; c < 3; c++) {
                float F = f + f0[l][c] * (1.0f - f);
                float Fr = (D * Vis) * F;
                float Fd = diffuse[l][c] * 0.3183098861837907f;
                color[l][c] = color[l][c] + (Fd + Fr) * L.color[c] * w;
            }
        }
    }
    for (int l = 0; l < 16; l++) {
        if (!active[l]) continue;
        for (int k = 0; k < 3; k++) color[l][k] = color[l][k] + base[l][k] * 0.05f;
    }
}

void lit_tile(const SurfaceLane* lanes, const bool* active, uint32_t x0, uint32_t y0, const float* tileDepth,
              const orc_resolve_params* p, float exposure, float color[16][3]) {
    float wpos[16][3] = {};
    for (int l = 0; l < 16; l++) {
        if (!active[l]) continue;
        // world position from the inverse screen projection (Shading.cpp:659-667)
        float px = (float)(int32_t)(x0 + (l & 3)), py = (float)(int32_t)(y0 + (l >> 2));
        float c[4];
        transform(p->inv_screen_proj, px, py, tileDepth[l], c);
        float rw = frcp(c[3]);
        wpos[l][0] = c[0] * rw; wpos[l][1] = c[1] * rw; wpos[l][2] = c[2] * rw;
    }
    lit_tile_at(lanes, active, wpos, p, exposure, color);
}

// RGBA8u::Pack of (r,g,b,a) (Texture.h:55-67): RNE, saturating packs.
uint32_t pack_rgba8(float r, float g, float b, float a) {
    float c[4] = { r, g, b, a };
    uint32_t out = 0;
    for (int k = 0; k < 4; k++) {
        int32_t i = cvt_rne(c[k] * 255.0f);
        i = std::min(std::max(i, -32768), 32767);  // packs_epi32
        i = std::min(std::max(i, 0), 255);          // packus_epi16
        out |= (uint32_t)i << (8 * k);
    }
    return out;
}

// ---------------------------------------------------------------- skybox

// MapOctahedron (Texture.h:282-288) followed by SampleLevel's scaling and
// rounding (Texture.h:412-416): 1/(|x| + |y| + |z|) from VRCP14PS alone,
// max(-z * w, 0) as vmaxps, (c * 0.5 + 0.5) contracted into one FMA.
void sky_coords(const float d[3], float slu, float slv, int32_t& ix, int32_t& iy) {
    const float w = rcp14((fabsf(d[0]) + fabsf(d[1])) + fabsf(d[2]));
    const float mz = -d[2] * w;
    const float t = mz > 0.0f ? mz : 0.0f;
    const float u = fmaf(fmaf(d[0], w, mulsign(t, d[0])), 0.5f, 0.5f);
    const float v = fmaf(fmaf(d[1], w, mulsign(t, d[1])), 0.5f, 0.5f);
    ix = cvt_rne(u * slu);
    iy = cvt_rne(v * slv);
}

// R11G11B10f::Unpack (Texture.h:157-182): exponent and mantissa moved into
// a float by bit operations (no denormal, NaN or infinity handling).
void unpack_r11g11b10(uint32_t x, float c[3]) {
    c[0] = bitsf((((x >> 21) << 17) & 0x0FFE0000u) + 0x38000000u);
    c[1] = bitsf((((x >> 10) << 17) & 0x0FFE0000u) + 0x38000000u);
    c[2] = bitsf(((x << 18) & 0x0FFC0000u) + 0x38000000u);
}

// SampleLinear's float path (Texture.h:557-571) on four texels: 8-bit
// fractions times 1/256, the x weight zeroed past the last column; the
// optimised build evaluates rowA + (rowB - rowA) * fy as
// fma(fy, fma(fx, c11 - c01, c01 - rowA), rowA).
void env_bilinear(const uint32_t d[4], int32_t ixf, int32_t iyf, bool inx, float out[3]) {
    float fx = (float)(ixf & 255) * (1.0f / 256);
    const float fy = (float)(iyf & 255) * (1.0f / 256);
    if (!inx) fx = 0.0f;
    float c[4][3];
    for (int k = 0; k < 4; k++) unpack_r11g11b10(d[k], c[k]);
    for (int ch = 0; ch < 3; ch++) {
        const float rowA = fmaf(fx, c[1][ch] - c[0][ch], c[0][ch]);
        out[ch] = fmaf(fy, fmaf(fx, c[3][ch] - c[2][ch], c[2][ch] - rowA), rowA);
    }
}

// SkyboxTex->SampleOctLevel<EnvSampler>(dir, 1) (Shading.cpp:677,
// Texture.h:467-480): level 1 (the mip fraction is 0), ClampToEdge, linear
// filter, layer 0 of an R11G11B10f texture in TiledY8 layout.
void sky_sample(const orc_texture* t, const float dir[3], float out[3]) {
    int32_t ix, iy;
    sky_coords(dir, t->scale_lerp_u, t->scale_lerp_v, ix, iy);
    ix = std::min(std::max(ix, 0), t->mask_lerp_u);
    iy = std::min(std::max(iy, 0), t->mask_lerp_v);
    const int32_t mip = std::min(1, (int32_t)t->mip_levels - 1);
    int32_t stride = (int32_t)t->row_shift;
    uint32_t offset = 0;
    if (mip > 0) {
        ix >>= mip;
        iy >>= mip;
        stride -= mip;
        offset = t->mip_offsets[mip];
    }
    const int32_t ixf = std::max(ix - 127, 0), iyf = std::max(iy - 127, 0);
    const int32_t tx = ixf >> 8, ty = iyf >> 8;
    const bool inx = ((tx + 1) << mip) < (int32_t)t->width;
    const bool iny = ((ty + 1) << mip) < (int32_t)t->height;
    const uint32_t o0 = offset + texel_offset(tx, ty, stride);
    const uint32_t o1 = offset + texel_offset(tx, ty + (iny ? 1 : 0), stride);
    const uint32_t d[4] = { t->data[o0], t->data[o0 + 8], t->data[o1], t->data[o1 + 8] };
    env_bilinear(d, ixf, iyf, inx, out);
}

// ---------------------------------------------------------------- light sprites

// cvttps2dq: truncation, out of range (and NaN) to INT32_MIN.
int32_t cvt_trunc(float x) {
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}

// Placement of a light's sprite (Shading.cpp:692-706) as the reference's
// optimised build computes it: glm's mat4 * vec4 becomes one FMA chain from
// the translation column, the divisions by w a multiply by 1/w (one IEEE
// division) and `/ 30` a multiply by 1/30.
struct Sprite {
    bool on;
    float depth, cx, cy, radius;
    int32_t x0, y0, x1, y1;        // tile range: x0, y0 multiples of 4
};

Sprite sprite_of(const float* m, const float pos[3], uint32_t W, uint32_t H) {
    Sprite s = {};
    float c[4];
    for (int r = 0; r < 4; r++) c[r] = fmaf(m[8 + r], pos[2], fmaf(m[4 + r], pos[1], fmaf(m[r], pos[0], m[12 + r])));
    float mx = fmaxf(fmaxf(fabsf(c[0]), fabsf(c[1])), fabsf(c[2]));
    if (mx > c[3]) return s;
    const float rw = 1.0f / c[3];
    s.depth = c[2] * rw;
    s.radius = ((float)std::max(W, H) * (1.0f / 30.0f)) * rw;
    s.cx = fmaf(c[0] * 0.5f, rw, 0.5f) * (float)W;
    s.cy = fmaf(c[1] * 0.5f, rw, 0.5f) * (float)H;
    s.x0 = std::max(cvt_trunc(s.cx - s.radius), 0) & ~3;
    s.y0 = std::max(cvt_trunc(s.cy - s.radius), 0) & ~3;
    s.x1 = std::min(cvt_trunc(s.cx + s.radius), (int32_t)W);
    s.y1 = std::min(cvt_trunc(s.cy + s.radius), (int32_t)H);
    s.on = true;
    return s;
}

uint32_t pack_byte(float v) {  // one channel of RGBA8u::Pack
    int32_t i = cvt_rne(v);
    i = std::min(std::max(i, -32768), 32767);
    return (uint32_t)std::min(std::max(i, 0), 255);
}

// Sprite colour of tile (x, y) (Shading.cpp:712-724): squared distance to the
// centre less r^2, a = 1 - (-d / r^2) evaluated as fma(d, 1/r^2, 1), alpha
// byte RNE(255 - 255 a^2), the light colour in the other bytes.
void sprite_tile(const Sprite& s, const float color[3], int32_t x, int32_t y, float* dist, uint32_t* fg) {
    const float hx = 0.5f - s.cx, hy = 0.5f - s.cy;
    const float r2 = s.radius * s.radius, inv = 1.0f / r2;
    const uint32_t rgb = pack_byte(color[0] * 255.0f) | (pack_byte(color[1] * 255.0f) << 8) |
                         (pack_byte(color[2] * 255.0f) << 16);
    for (int l = 0; l < 16; l++) {
        const float rx = (float)(x + (l & 3)) + hx, ry = (float)(y + (l >> 2)) + hy;
        const float d = fmaf(rx, rx, ry * ry) - r2;
        const float a = fmaf(d, inv, 1.0f);
        dist[l] = d;
        fg[l] = rgb | (pack_byte(fmaf(-(a * a), 255.0f, 255.0f)) << 24);
    }
}

// AlphaBlendU8 (Shading.cpp:238-246): source over in 8.8 fixed point.
uint32_t alpha_blend_u8(uint32_t bg, uint32_t fg) {
    uint32_t t = (fg >> 1) & 0x7F800000u;
    t |= (t >> 16) | 0x00400040u;
    uint32_t rb = lerp16(bg & 0x00FF00FFu, fg & 0x00FF00FFu, t);
    uint32_t ag = lerp16((bg >> 8) & 0x00FF00FFu, (fg >> 8) & 0x00FF00FFu, t);
    return rb | (ag << 8);
}

// The light sprites Resolve draws over its output (Shading.cpp:689-727):
// for each point or spot light in front of the camera, every 4x4 tile of the
// sprite's range with a pixel inside the disc and behind the light blends
// the light colour over the pixels behind the light.
void draw_light_sprites(uint32_t* color, const float* depth, uint32_t W, uint32_t H, const orc_resolve_params* p,
                        int32_t row0, int32_t row1) {
    for (uint32_t li = 0; li < p->num_lights; li++) {
        const orc_light& L = p->lights[li];
        if (L.type == 0) continue;
        const Sprite s = sprite_of(p->world_to_clip, L.position, W, H);
        if (!s.on) continue;
        for (int32_t y = std::max(s.y0, row0); y < std::min(s.y1, row1); y += 4) {
            for (int32_t x = s.x0; x < s.x1; x += 4) {
                const uint32_t off = pixel_offset(W, (uint32_t)x, (uint32_t)y);
                float d[16];
                uint32_t fg[16];
                sprite_tile(s, L.color, x, y, d, fg);
                bool any = false;
                for (int l = 0; l < 16; l++) any |= d[l] < 0.0f && s.depth > depth[off + l];
                if (!any) continue;
                for (int l = 0; l < 16; l++)
                    if (s.depth > depth[off + l]) color[off + l] = alpha_blend_u8(color[off + l], fg[l]);
            }
        }
    }
}

}  // namespace

// ====================================================================== API

extern "C" {

void orc_load_tables(const uint16_t* tables) {
    memcpy(g_rcp, tables, sizeof(g_rcp));
    memcpy(g_rsq, tables + 65536, sizeof(g_rsq));
}
float orc_rcp14(float x) { return rcp14(x); }
void orc_approx_n(int32_t op, const float* a, const float* b, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
        out[i] = op == 0 ? rcp14(a[i]) : op == 1 ? rsqrt14(a[i]) : fdiv(a[i], b[i]);
}
float orc_rsqrt14(float x) { return rsqrt14(x); }
float orc_div(float a, float b) { return fdiv(a, b); }
int32_t orc_cvt_rne(float x) { return cvt_rne(x); }

void orc_fb_clear(orc_framebuffer* fb, uint32_t color, float depth) {  // Rasterizer.h:35-48
    uint32_t n = fb->width * fb->height;
    for (uint32_t i = 0; i < n; i++) fb->data[i] = color;
    uint32_t db = fbits(depth);
    for (uint32_t i = 0; i < n; i++) fb->data[fb->layer_stride + i] = db;
}

void orc_get_pixels(const orc_framebuffer* fb, uint32_t layer, uint32_t* dest, uint32_t stride) {
    const uint32_t* src = fb->data + (size_t)layer * fb->layer_stride;
    for (uint32_t y = 0; y < fb->height; y++)
        for (uint32_t x = 0; x < fb->width; x++)
            dest[(size_t)y * stride + x] = src[pixel_offset(fb->width, x, y)];
}

void orc_draw_meshlets(orc_framebuffer* fb, uint32_t count, const orc_shading_ctx* ctx,
                       const orc_raster_opts* opts, orc_counters* counters) {
    orc_draw_meshlets_rows(fb, count, ctx, opts, counters, 0, fb->height);
}

void orc_draw_meshlets_rows(orc_framebuffer* fb, uint32_t count, const orc_shading_ctx* ctx,
                            const orc_raster_opts* opts, orc_counters* counters, uint32_t row0, uint32_t row1) {
    orc_counters localCtr = {};
    orc_counters* ctr = counters ? counters : &localCtr;
    DrawState st = { fb, opts, ctr, (int)row0, (int)row1 };
    int hw = (int)fb->width / 2, hh = (int)fb->height / 2;
    bool guard = !opts || opts->enable_guardband;
    bool clipping = !opts || opts->enable_clipping;
    float bx = guard ? (float)kMaxRenderSize / (float)fb->width : 1.0f;
    float by = guard ? (float)kMaxRenderSize / (float)fb->height : 1.0f;

    ShadedMeshlet mesh;
    for (uint32_t meshIdx = 0; meshIdx < count; meshIdx++) {
        mesh = ShadedMeshlet();
        memset(mesh.Position, 0, sizeof(mesh.Position));
        shade_meshlet(*ctx, meshIdx, mesh);
        if (mesh.PrimCount == 0) continue;
        ctr->triangles_processed += mesh.PrimCount;
        uint32_t meshletId = ctx->meshlet_offset + meshIdx;
        // fragment shader inputs (FS_EncodeSurfaceId<AlphaTest>)
        FragmentCtx fc = {};
        const uint32_t matId = ctx->meshlets->material_id[meshletId];
        const bool deferred = opts && opts->pipeline == ORC_PIPELINE_DEFERRED;
        fc.material = matId != UINT32_MAX;
        fc.object_to_world = ctx->object_to_world;
        if (deferred && matId != UINT32_MAX && ctx->materials && ctx->materials[matId].texture >= 0) {
            // FS_EncodeGBuffer samples the base colour of every material
            // (Shading.cpp:357-366); a material without a texture would be
            // sampled through a null pointer there and is drawn here with
            // depth only
            fc.tex = &ctx->textures[ctx->materials[matId].texture];
            fc.cutoff = ctx->materials[matId].alpha_cutoff;
        } else if (!deferred && mesh.FragmentShaderId == 1 && matId != UINT32_MAX && ctx->materials &&
                   ctx->materials[matId].texture >= 0) {
            fc.alpha_test = true;
            fc.tex = &ctx->textures[ctx->materials[matId].texture];
            fc.cutoff = ctx->materials[matId].alpha_cutoff;
        }
        auto fragment_ctx = [&](uint32_t prim) {
            FragmentCtx f = fc;
            if (f.alpha_test || f.tex) {
                for (int k = 0; k < 3; k++) {
                    // indices of a well-formed meshlet are below 64; the mask
                    // keeps a malformed one inside the meshlet's 64 slots
                    uint32_t vi = mesh.Indices[k][prim] & 63u;
                    uint32_t tc = ctx->meshlets->texcoords[(size_t)meshletId * 64 + vi];
                    f.uv[k][0] = half_to_float((uint16_t)(tc & 0xFFFF));
                    f.uv[k][1] = half_to_float((uint16_t)(tc >> 16));
                    f.nt[k] = ctx->meshlets->normal_tangents[(size_t)meshletId * 64 + vi];
                }
                const uint32_t v0 = mesh.Indices[0][prim] & 63u;
                f.handed = (uint32_t)((ctx->meshlets->tangent_handedness[meshletId] >> v0) & 1u) << 31;
            }
            return f;
        };

        for (uint32_t primOffset = 0; primOffset < mesh.PrimCount; primOffset += 16) {
            V4 v[3][16];
            for (int k = 0; k < 3; k++) {
                for (int l = 0; l < 16; l++) {
                    // gather_preload64 (SIMD.h:219-227): two 32-entry
                    // permutes chosen by idx < 32
                    uint32_t ix = mesh.Indices[k][primOffset + l];
                    uint32_t vi = ix < 32 ? ix : 32 + (ix & 31);
                    v[k][l] = { mesh.Position[0][vi], mesh.Position[1][vi], mesh.Position[2][vi], mesh.Position[3][vi] };
                }
            }
            ClipCodes cc = compute_clip_codes(v, bx, by);
            if (primOffset + 16 > mesh.PrimCount) {
                uint16_t trail = (uint16_t)((1u << (mesh.PrimCount % 16)) - 1);
                cc.AcceptMask &= trail;
                cc.NonTrivialMask &= trail;
            }
            if (cc.AcceptMask != 0) {
                TrianglePacket tris;
                uint16_t accept = setup_packet(tris, v, hw, hh, mesh.CullMode) & cc.AcceptMask;
                if (accept) {
                    EdgeVars e;
                    setup_edges(e, tris, hw, hh);
                    for (int l = 0; l < 16; l++) {
                        if (!((accept >> l) & 1)) continue;
                        uint32_t bmin, bmax;
                        render_bbox(tris, l, hw, hh, bmin, bmax);
                        draw_triangle(st, e, l, bmin, bmax, meshletId * 128 + primOffset + (uint32_t)l,
                                      fragment_ctx(primOffset + (uint32_t)l));
                        ctr->triangles_rasterized++;
                    }
                }
            }
            if (cc.NonTrivialMask != 0 && clipping) {
                for (int l = 0; l < 16; l++) ctr->triangles_clipped += (cc.NonTrivialMask >> l) & 1;
                static thread_local Clipper clipper;
                clipper.bx = bx;
                clipper.by = by;
                clipper.OutIdx = 0;
                clipper.clip(v, cc, hw, hh, mesh.CullMode);
                for (uint32_t i = 0; i < clipper.OutIdx; i += 16) {
                    const TrianglePacket& tris = clipper.Tris[i / 16];
                    EdgeVars e;
                    setup_edges(e, tris, hw, hh);
                    uint16_t drawMask = clipper.Masks[i / 16];
                    for (int l = 0; l < 16; l++) {
                        if (!((drawMask >> l) & 1)) continue;
                        uint32_t bmin, bmax;
                        render_bbox(tris, l, hw, hh, bmin, bmax);
                        uint32_t prim = clipper.PrimId[i / 16][l];
                        FragmentCtx f = fragment_ctx(primOffset + prim);
                        f.clipped = true;
                        for (int k = 0; k < 3; k++) {
                            f.ru[k] = clipper.ClipU[i / 16][l][k];
                            f.rv[k] = clipper.ClipV[i / 16][l][k];
                        }
                        draw_triangle(st, e, l, bmin, bmax, meshletId * 128 + primOffset + prim, f);
                        ctr->triangles_rasterized++;
                    }
                }
            }
        }
    }
}

// CullMeshlets frustum branch (Shading.cpp:797-809, 865-866); the planes are
// computed by the caller (Gribb-Hartmann, Shading.cpp:783-791).
uint32_t orc_cull_meshlets(uint16_t* bitmap, const orc_meshlets* m, uint32_t first, uint32_t count,
                           const float planes[20]) {
    return orc_cull_meshlets_occlusion(bitmap, m, first, count, planes, nullptr);
}

uint32_t orc_hiz_levels(uint32_t width, uint32_t height) {
    // CreateFramebuffer's depth map (Main.cpp:54-56): power-of-two sides
    // above (size - 1) / 2, CreateTexture2D's level rule (Texture.h:606-611)
    auto half = [](uint32_t v) { uint32_t h = (v - 1) / 2; return h ? 1u << (32 - __builtin_clz(h)) : 1u; };
    uint32_t hw = half(width), hh = half(height), mip = 0;
    for (; mip < 16; mip++)
        if ((hw >> mip) < 4 || (hh >> mip) < 4) break;
    return mip;
}

void orc_hiz_build(const orc_framebuffer* fb, float* out) {
    // texutil::DownsampleDepth (ImageHelpers.cpp:149-247): level m texel (i, j)
    // is the minimum depth over pixels [i, i+1) x [j, j+1) * 2^(m+1) inside
    // the framebuffer (pixels outside count as FLT_MAX). The depth map's
    // top level is never sampled, so levels 0 .. orc_hiz_levels - 2 are kept.
    const uint32_t W = fb->width, H = fb->height;
    const float* depth = (const float*)(fb->data + fb->layer_stride);
    const uint32_t L = orc_hiz_levels(W, H);
    size_t off = 0;
    for (uint32_t m = 0; m + 1 < L; m++) {
        const uint32_t s = 1u << (m + 1);
        const uint32_t w = ((W - 1) >> (m + 1)) + 1, h = ((H - 1) >> (m + 1)) + 1;
        for (uint32_t j = 0; j < h; j++)
            for (uint32_t i = 0; i < w; i++) {
                float v = FLT_MAX;
                for (uint32_t y = j * s; y < std::min((j + 1) * s, H); y++)
                    for (uint32_t x = i * s; x < std::min((i + 1) * s, W); x++)
                        v = std::min(v, depth[pixel_offset(W, x, y)]);
                out[off + (size_t)j * w + i] = v;
            }
        off += (size_t)w * h;
    }
}

// ProjectSphere (Shading.cpp:264-279) and the occlusion test of
// CullMeshlets (Shading.cpp:811-840) for one bound sphere, in the operation
// order of the reference's optimised build. Returns whether the meshlet
// stays visible.
static void hiz_sphere(const float* b, const orc_occlusion* o, bool& proj, float bb[4], int32_t& lv, int32_t rect[4],
                float& dsphere) {
    const float* m = o->object_to_prev_view;
    const float cx = fmaf(b[0], m[0], fmaf(b[1], m[4], fmaf(b[2], m[8], m[12])));
    const float cy = fmaf(b[0], m[1], fmaf(b[1], m[5], fmaf(b[2], m[9], m[13])));
    const float cz = fmaf(b[0], m[2], fmaf(b[1], m[6], fmaf(b[2], m[10], m[14])));
    const float r = b[3] * o->object_to_world_scale;
    const float nz = -cz;
    proj = (r + o->proj[0]) <= nz;
    const float crx = cx * r, cry = cy * r, crz = r * nz;
    const float ax = fmaf(cx, cx, -(r * r)) + cz * cz, ay = fmaf(cy, cy, -(r * r)) + cz * cz;
    const float vx = ax * rsqrt14(ax), vy = ay * rsqrt14(ay);
    const float sxh = o->proj[1] * 0.5f, syh = o->proj[2] * 0.5f;
    bb[0] = fmaf((vx * cx - crz) * rcp14(vx * nz + crx), sxh, 0.5f);
    bb[1] = fmaf((vy * cy + crz) * rcp14(vy * nz - cry), syh, 0.5f);
    bb[2] = fmaf((vx * cx + crz) * rcp14(vx * nz - crx), sxh, 0.5f);
    bb[3] = fmaf((vy * cy - crz) * rcp14(vy * nz + cry), syh, 0.5f);
    const float fw = o->frame_size[0], fh = o->frame_size[1];
    const float sx = (bb[2] - bb[0]) * fw, sy = (bb[3] - bb[1]) * fh;
    const float mx = (sx > sy) ? sx : sy;           // vmaxps: second operand on NaN
    lv = wrap_sub((int32_t)fbits(mx), 127 << 23) >> 23;
    lv = std::min(std::max(lv, 1), (int32_t)o->mip_levels - 1);
    rect[0] = std::max(cvt_trunc(bb[0] * fw), 0) >> lv;
    rect[1] = std::max(cvt_trunc(bb[1] * fh), 0) >> lv;
    rect[2] = std::min(cvt_trunc(bb[2] * fw), cvt_trunc(fw) - 1) >> lv;
    rect[3] = std::min(cvt_trunc(bb[3] * fh), cvt_trunc(fh) - 1) >> lv;
    dsphere = fdiv(o->proj[0], -(cz + r));
}

static bool hiz_visible(const float* b, const orc_occlusion* o) {
    bool proj;
    float bb[4], dsphere;
    int32_t lv, rect[4];
    hiz_sphere(b, o, proj, bb, lv, rect, dsphere);
    if (!proj) return true;
    // the depth map's level lv - 1 (its texels cover 2^lv pixels)
    const uint32_t W = (uint32_t)o->frame_size[0], H = (uint32_t)o->frame_size[1];
    size_t off = 0;
    for (int32_t k = 0; k < lv - 1; k++) off += (size_t)(((W - 1) >> (k + 1)) + 1) * (((H - 1) >> (k + 1)) + 1);
    const uint32_t w = ((W - 1) >> lv) + 1;
    float vis = FLT_MAX;
    for (int32_t y = rect[1]; y <= rect[3]; y++)
        for (int32_t x = rect[0]; x <= rect[2]; x++) vis = std::min(vis, o->depth_map[off + (size_t)y * w + x]);
    return dsphere > vis;
}

uint32_t orc_cull_meshlets_occlusion(uint16_t* bitmap, const orc_meshlets* m, uint32_t first, uint32_t count,
                                     const float planes[20], const orc_occlusion* occ) {
    uint32_t visible = 0;
    for (uint32_t off = 0; off < count; off += 16) {
        uint16_t mask = 0;
        for (uint32_t l = 0; l < 16 && off + l < count; l++) {
            const float* b = m->bounds + (size_t)(first + off + l) * 4;
            bool vis = true;
            for (int i = 0; i < 5; i++) {
                const float* P = planes + i * 4;
                float d = fmaf(b[0], P[0], fmaf(b[1], P[1], b[2] * P[2])) + P[3];
                vis = vis && (d > -b[3]);
            }
            if (vis && occ) vis = hiz_visible(b, occ);
            if (vis) { mask |= (uint16_t)(1u << l); visible++; }
        }
        bitmap[off / 16] = mask;
    }
    return visible;
}

void orc_hiz_sphere16(const float* sphere, const orc_occlusion* o, int32_t* mask, float* bb, int32_t* mip,
                      int32_t* rect, float* depth) {
    for (int l = 0; l < 16; l++) {
        const float b[4] = { sphere[l], sphere[16 + l], sphere[32 + l], sphere[48 + l] };
        bool proj;
        float q[4];
        int32_t r[4];
        hiz_sphere(b, o, proj, q, mip[l], r, depth[l]);
        mask[l] = proj ? -1 : 0;
        for (int k = 0; k < 4; k++) {
            bb[k * 16 + l] = q[k];
            rect[k * 16 + l] = r[k];
        }
    }
}

void orc_sample_level16(const orc_texture* tex, const float* u, const float* v, int32_t layer,
                        const int32_t* mip, int32_t wrap, int32_t mag_linear, int32_t min_linear,
                        uint32_t* out) {
    sample_level16(tex, u, v, layer, mip, wrap, mag_linear != 0, min_linear != 0, nullptr, out);
}

void orc_intersect_triangle16(const float* clip, const float* suv, float sw, float sh,
                              float* bary, float* ddx, float* ddy) {
    for (int l = 0; l < 16; l++) {
        float c[3][4];
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 4; j++) c[k][j] = clip[(k * 4 + j) * 16 + l];
        float b[3], dx[3], dy[3];
        intersect_triangle(c, suv[l], suv[16 + l], sw, sh, b, dx, dy);
        for (int k = 0; k < 3; k++) {
            bary[k * 16 + l] = b[k];
            ddx[k * 16 + l] = dx[k];
            ddy[k * 16 + l] = dy[k];
        }
    }
}

void orc_setup16(const float* clip, int32_t w, int32_t h, int32_t cull_mode,
                 int32_t* X, int32_t* Y, float* Z, float* W, float* det, uint16_t* mask) {
    V4 v[3][16];
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 16; l++)
            v[k][l] = { clip[(k * 4 + 0) * 16 + l], clip[(k * 4 + 1) * 16 + l], clip[(k * 4 + 2) * 16 + l], clip[(k * 4 + 3) * 16 + l] };
    TrianglePacket t;
    *mask = setup_packet(t, v, w / 2, h / 2, cull_mode);
    uint32_t* P[3] = { t.Pos0, t.Pos1, t.Pos2 };
    float* Zs[3] = { t.Z0, t.Z1, t.Z2 };
    float* Ws[3] = { t.W0, t.W1, t.W2 };
    for (int k = 0; k < 3; k++) {
        for (int l = 0; l < 16; l++) {
            X[k * 16 + l] = lo16(P[k][l]);
            Y[k * 16 + l] = (int32_t)P[k][l] >> 16;
            Z[k * 16 + l] = Zs[k][l];
            W[k * 16 + l] = Ws[k][l];
        }
    }
    // det is not kept by the packet; recompute it for the probe comparison
    for (int l = 0; l < 16; l++) {
        float x[3], y[3];
        for (int k = 0; k < 3; k++) {
            float rw = frcp(v[k][l].w);
            x[k] = v[k][l].x * rw;
            y[k] = v[k][l].y * rw;
        }
        float d = fmaf(x[2] - x[0], y[1] - y[0], (x[0] - x[1]) * (y[2] - y[0]));
        if (cull_mode != 1) {
            bool flip = (cull_mode == 2) ? true : (d < 0);
            d = flip ? -d : d;
        }
        det[l] = d;
    }
}

// Float part of TriangleEdgeVars::Setup for 16 lanes (Rasterizer.cpp:320-323).
void orc_edge16(const int32_t* det, const float* Z, float* Z10, float* Z20) {
    for (int l = 0; l < 16; l++) {
        float rcpArea = fdiv(16.0f, (float)det[l]);
        Z10[l] = (Z[16 + l] - Z[l]) * rcpArea;
        Z20[l] = (Z[32 + l] - Z[l]) * rcpArea;
    }
}

// Texture coordinates, UV gradients and mip level for 16 lanes, the
// arithmetic of resolve_surface_tile (Shading.cpp:516-539).
void orc_uvgrad16(const float* t, const float* bary, const float* ddx, const float* ddy, float su, float sv,
                  float* uv, int32_t* mip) {
    for (int l = 0; l < 16; l++) {
        float u0 = t[l], v0 = t[16 + l];
        float t10x = t[32 + l] - u0, t10y = t[48 + l] - v0;
        float t20x = t[64 + l] - u0, t20y = t[80 + l] - v0;
        float by = bary[16 + l], bz = bary[32 + l];
        uv[l] = fmaf(bz, t20x, fmaf(by, t10x, u0));
        uv[16 + l] = fmaf(bz, t20y, fmaf(by, t10y, v0));
        float gx = fmaf(t20x, ddx[32 + l], t10x * ddx[16 + l]);
        float gy = fmaf(t20y, ddx[32 + l], t10y * ddx[16 + l]);
        float gz = fmaf(t20x, ddy[32 + l], t10x * ddy[16 + l]);
        float gw = fmaf(t20y, ddy[32 + l], t10y * ddy[16 + l]);
        float dx = fmaf(gx, gx, gy * gy) * (su * su);
        float dy = fmaf(gz, gz, gw * gw) * (sv * sv);
        float mx = (dx > dy) ? dx : dy;
        mip[l] = wrap_sub((int32_t)fbits(mx), 127 << 23) >> 24;
    }
}

// Clipper::ClipTriangles for a single triangle (lane 0 of a packet): returns
// the fan triangles' clip-space vertices [n][3][4] and their count.
int32_t orc_clip_triangle(const float* v /*[3][4]*/, uint32_t outcodes, float gbx, float gby, float* out) {
    V4 vin[3][16];
    memset(vin, 0, sizeof(vin));
    for (int k = 0; k < 3; k++) vin[k][0] = { v[k * 4 + 0], v[k * 4 + 1], v[k * 4 + 2], v[k * 4 + 3] };
    ClipCodes cc;
    cc.NonTrivialMask = 1;
    cc.OutCodes[0] = (uint8_t)outcodes;
    Clipper c;
    c.bx = gbx;
    c.by = gby;
    c.OutIdx = 0;
    c.clip(vin, cc, 1 << 14, 1 << 14, 0);  // large viewport: only the geometry is used
    for (uint32_t i = 0; i < c.OutIdx; i++) {
        for (int k = 0; k < 3; k++) {
            const V4& q = c.Out[k][i / 16][i % 16];
            out[(i * 3 + k) * 4 + 0] = q.x;
            out[(i * 3 + k) * 4 + 1] = q.y;
            out[(i * 3 + k) * 4 + 2] = q.z;
            out[(i * 3 + k) * 4 + 3] = q.w;
        }
    }
    return (int32_t)c.OutIdx;
}

// Coefficients of ColormapTurbo (Shading.cpp:249-260): Google's published
// degree-5 polynomial fit of the Turbo colormap, one row per channel,
// constant term first; double literals rounded to float.
static const float kTurbo[18] = {
    0.13572138, 4.61539260,  -42.66032258, 132.13108234, -152.94239396, 59.28637943,
    0.09140261, 2.19418839,  4.84296658,   -14.18503333, 4.27729857,    2.82956604,
    0.10667330, 12.64194608, -60.58204836, 110.36276771, -89.90310912,  27.34824973,
};

// The overdraw debug colour (ResolveDebug, Shading.cpp:759-766) of one
// counter word (pixel count << 16 | quad-helper count): ColormapTurbo of
// pixels / 30 (OverdrawPixel) or (pixels + helpers * 0.5) / 30
// (OverdrawQuad). The fast-math build divides by multiplying with 1/30 and
// folds the innermost x * c5 into s * (c5 * (1/30)), s the undivided count;
// the rest of the Horner chain is FMAs.
static void overdraw_colour(uint32_t data, bool quad, float rgb[3]) {
    const float pix = (float)(data >> 16);
    const float s = quad ? pix + (float)(data & 0xFFFFu) * 0.5f : pix;
    const float r = 1.0f / 30.0f;
    const float x = s * r;
    for (int k = 0; k < 3; k++) {
        const float* c = kTurbo + 6 * k;
        float t = fmaf(s, c[5] * r, c[4]);
        t = fmaf(x, t, c[3]);
        t = fmaf(x, t, c[2]);
        t = fmaf(x, t, c[1]);
        rgb[k] = fmaf(x, t, c[0]);
    }
}

// ResolveDebug(DebugLayer::BaseColor) (Shading.cpp:734-773).
// ResolveDebug (Shading.cpp:734-773) for the surface layers (BaseColor,
// Normals, MetallicRoughness) and the id layers (MeshletId, TriangleId);
// uncovered pixels get the background checker.
void orc_resolve_debug(orc_framebuffer* fb, const orc_shading_ctx* ctx, uint32_t layer, const float* object_to_world) {
    uint32_t W = fb->width, H = fb->height;
    float sw = (float)W, sh = (float)H;
    float scaleX = 2.0f / sw, scaleY = 2.0f / sh;
    float centerX = fmaf(0.5f, scaleX, -1.0f), centerY = fmaf(0.5f, scaleY, -1.0f);
    const float* depth = (const float*)(fb->data + fb->layer_stride);
    std::vector<uint32_t> out(fb->layer_stride);
    memcpy(out.data(), fb->data, (size_t)W * H * 4);
    const bool surface = layer <= ORC_DEBUG_METALLIC_ROUGHNESS;
    const float* o2w = (layer == ORC_DEBUG_NORMALS || layer == ORC_DEBUG_METALLIC_ROUGHNESS) ? object_to_world : nullptr;
    for (uint32_t y0 = 0; y0 < H; y0 += 4) {
        for (uint32_t x0 = 0; x0 < W; x0 += 4) {
            uint32_t off = pixel_offset(W, x0, y0);
            bool active[16];
            float sx[16], sy[16];
            uint32_t ids[16];
            for (int l = 0; l < 16; l++) {
                active[l] = !(depth[off + l] <= 0.0f);
                ids[l] = fb->data[off + l];
                sx[l] = fmaf((float)(int32_t)(x0 + (l & 3)), scaleX, centerX);
                sy[l] = fmaf((float)(int32_t)(y0 + (l >> 2)), scaleY, centerY);
            }
            SurfaceLane lanes[16];
            bool anyActive = false;
            for (int l = 0; l < 16; l++) anyActive |= active[l];
            if (anyActive && surface) resolve_surface_tile(*ctx, ids, active, sx, sy, sw, sh, lanes, o2w);
            uint32_t bg = ((x0 ^ y0) & 4) ? 0xFFA0A0A0u : 0xFFFFFFFFu;
            for (int l = 0; l < 16; l++) {
                if (!active[l]) { out[off + l] = bg; continue; }
                float c[4];
                if (layer == ORC_DEBUG_BASE_COLOR) {
                    unpack_rgba8(lanes[l].albedo, c);
                } else if (layer == ORC_DEBUG_NORMALS) {
                    // Normal * 0.5 + 0.5 (one FMA in the fast-math build; x * 0.5 is exact)
                    for (int k = 0; k < 3; k++) c[k] = fmaf(lanes[l].nrm[k], 0.5f, 0.5f);
                } else if (layer == ORC_DEBUG_METALLIC_ROUGHNESS) {
                    c[0] = (float)((lanes[l].nmr >> 16) & 255) * (1.0f / 255);
                    c[1] = (float)((lanes[l].nmr >> 24) & 255) * (1.0f / 255);
                    c[2] = 0.0f;
                } else if (layer == ORC_DEBUG_OVERDRAW_PIXEL || layer == ORC_DEBUG_OVERDRAW_QUAD) {
                    overdraw_colour(ids[l], layer == ORC_DEBUG_OVERDRAW_QUAD, c);
                } else {
                    // RGBA8u::Unpack of the hashed id (MeshletId: id / 128)
                    uint32_t v = layer == ORC_DEBUG_MESHLET_ID ? ids[l] / 128u : ids[l];
                    unpack_rgba8(v * 123456789u, c);
                }
                out[off + l] = pack_rgba8(c[0], c[1], c[2], 1.0f);
            }
        }
    }
    memcpy(fb->data, out.data(), (size_t)W * H * 4);
}

void orc_sky_coords16(const float* dir, float slu, float slv, int32_t* ix, int32_t* iy) {
    for (int l = 0; l < 16; l++) {
        const float d[3] = { dir[l], dir[16 + l], dir[32 + l] };
        sky_coords(d, slu, slv, ix[l], iy[l]);
    }
}

void orc_env_bilinear16(const uint32_t* d, const int32_t* ixf, const int32_t* iyf, const int32_t* inx, float* out) {
    for (int l = 0; l < 16; l++) {
        const uint32_t t[4] = { d[l], d[16 + l], d[32 + l], d[48 + l] };
        float c[3];
        env_bilinear(t, ixf[l], iyf[l], inx[l] != 0, c);
        for (int k = 0; k < 3; k++) out[16 * k + l] = c[k];
    }
}

void orc_gbuffer_pack16(const float* n, const uint32_t* nmr, uint32_t* out) {
    for (int l = 0; l < 16; l++) {
        const float v[3] = { n[l], n[16 + l], n[32 + l] };
        out[l] = gbuffer_pack(v, nmr[l]);
    }
}

void orc_overdraw_colour16(const uint32_t* data, int32_t quad, float* out) {
    for (int l = 0; l < 16; l++) {
        float c[3];
        overdraw_colour(data[l], quad != 0, c);
        for (int k = 0; k < 3; k++) out[k * 16 + l] = c[k];
    }
}

void orc_resolve_albedo(orc_framebuffer* fb, const orc_shading_ctx* ctx) {
    orc_resolve_debug(fb, ctx, ORC_DEBUG_BASE_COLOR, nullptr);
}

// Resolve (Shading.cpp:658-727) without the skybox: covered pixels get
// EvalLighting of their surface, uncovered pixels are black; then the light
// sprites when asked for.
void orc_resolve_lit(orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_resolve_params* p) {
    orc_resolve_lit_rows(fb, ctx, p, 0, fb->height);
}

void orc_resolve_lit_rows(orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_resolve_params* p, uint32_t row0,
                          uint32_t row1) {
    uint32_t W = fb->width, H = fb->height;
    float sw = (float)W, sh = (float)H;
    float scaleX = 2.0f / sw, scaleY = 2.0f / sh;
    float centerX = fmaf(0.5f, scaleX, -1.0f), centerY = fmaf(0.5f, scaleY, -1.0f);
    const float* depth = (const float*)(fb->data + fb->layer_stride);
    std::vector<uint32_t> out((size_t)W * H);
    const float exposure = p->exposure;
    const float lightExposure = p->exposure * 0.001f;
    for (uint32_t y0 = row0; y0 < row1; y0 += 4) {
        for (uint32_t x0 = 0; x0 < W; x0 += 4) {
            uint32_t off = pixel_offset(W, x0, y0);
            bool active[16];
            float sx[16], sy[16];
            uint32_t ids[16];
            bool anyActive = false;
            for (int l = 0; l < 16; l++) {
                active[l] = !(depth[off + l] <= 0.0f);
                anyActive |= active[l];
                ids[l] = fb->data[off + l];
                sx[l] = fmaf((float)(int32_t)(x0 + (l & 3)), scaleX, centerX);
                sy[l] = fmaf((float)(int32_t)(y0 + (l >> 2)), scaleY, centerY);
            }
            float color[16][3] = {};
            if (anyActive) {
                SurfaceLane lanes[16];
                resolve_surface_tile(*ctx, ids, active, sx, sy, sw, sh, lanes, p->object_to_world);
                lit_tile(lanes, active, x0, y0, depth + off, p, lightExposure, color);
            }
            for (int l = 0; l < 16; l++) {
                if (active[l] || !p->skybox) continue;
                // sky: the view direction through the pixel at depth 1
                // (Shading.cpp:663-667, 675-678)
                float px = (float)(int32_t)(x0 + (l & 3)), py = (float)(int32_t)(y0 + (l >> 2));
                float c4[4];
                transform(p->inv_screen_proj, px, py, 1.0f, c4);
                const float rw = frcp(c4[3]);
                const float dir[3] = { c4[0] * rw - p->view_pos[0], c4[1] * rw - p->view_pos[1],
                                       c4[2] * rw - p->view_pos[2] };
                sky_sample(p->skybox, dir, color[l]);
            }
            for (int l = 0; l < 16; l++) {
                float c[3];
                for (int k = 0; k < 3; k++) {
                    // Tonemap_Unreal(x * Exposure) (Shading.cpp:221-226, 680)
                    float x = (active[l] || p->skybox) ? color[l][k] * exposure : 0.0f;
                    c[k] = fdiv(x, x + 0.155f) * 1.019f;
                }
                out[(size_t)off + l] = pack_rgba8(c[0], c[1], c[2], 1.0f);
            }
        }
    }
    if (p->light_sprites) draw_light_sprites(out.data(), depth, W, fb->height, p, (int32_t)row0, (int32_t)row1);
    // rows [row0, row1) of 4x4 tiles are one contiguous range of the layer
    memcpy(fb->data + (size_t)row0 * W, out.data() + (size_t)row0 * W, (size_t)(row1 - row0) * W * 4);
}

void orc_eval_lighting16(const uint32_t* albedo, const uint32_t* nmr, const float* normal, const float* wpos,
                         const uint8_t* active, const orc_resolve_params* p, uint32_t* out) {
    SurfaceLane lanes[16] = {};
    bool act[16];
    float wp[16][3];
    for (int l = 0; l < 16; l++) {
        act[l] = active[l] != 0;
        lanes[l].albedo = albedo[l];
        lanes[l].nmr = nmr[l];
        for (int k = 0; k < 3; k++) {
            lanes[l].nrm[k] = normal[k * 16 + l];
            wp[l][k] = wpos[k * 16 + l];
        }
    }
    float color[16][3] = {};
    lit_tile_at(lanes, act, wp, p, p->exposure * 0.001f, color);
    for (int l = 0; l < 16; l++) {
        float c[3];
        for (int k = 0; k < 3; k++) {
            // Tonemap_Unreal(x * Exposure) (Shading.cpp:221-226, 680)
            float x = act[l] ? color[l][k] * p->exposure : 0.0f;
            c[k] = fdiv(x, x + 0.155f) * 1.019f;
        }
        out[l] = pack_rgba8(c[0], c[1], c[2], 1.0f);
    }
}

void orc_unpack_srgb_n(const uint32_t* packed, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) unpack_srgb(packed[i], out + 4 * i);
}

// ---------------------------------------------------------------- differentiable resolve

void orc_fragment16(const int32_t* e1, const int32_t* e2, const float* coef, float* depth, float* bu, float* bv) {
    for (int l = 0; l < 16; l++) {
        float u = (float)e1[l], v = (float)e2[l];
        depth[l] = fmaf(u, coef[1], fmaf(v, coef[2], coef[0]));
        fragment_bary(u, v, coef[3], coef[4], coef[5], coef[6], bu[l], bv[l]);
    }
}

void orc_normal_map16(const float* n, const float* t, const uint32_t* hand, const uint32_t* nmr, float* out) {
    for (int l = 0; l < 16; l++) {
        float nn[3] = { n[l], n[16 + l], n[32 + l] }, tt[3] = { t[l], t[16 + l], t[32 + l] }, o[3];
        normal_map(nn, tt, hand[l], nmr[l], o);
        for (int k = 0; k < 3; k++) out[k * 16 + l] = o[k];
    }
}

int32_t orc_sprite_place(const float* m, const float* pos, uint32_t W, uint32_t H, float* out, int32_t* tiles) {
    const Sprite s = sprite_of(m, pos, W, H);
    if (!s.on) return 0;
    out[0] = s.depth; out[1] = s.cx; out[2] = s.cy; out[3] = s.radius;
    tiles[0] = s.x0; tiles[1] = s.y0; tiles[2] = s.x1; tiles[3] = s.y1;
    return 1;
}

void orc_sprite_tile(int32_t x, int32_t y, float cx, float cy, float radius, const float* color, float* dist,
                     uint32_t* fg) {
    Sprite s = {};
    s.cx = cx; s.cy = cy; s.radius = radius;
    sprite_tile(s, color, x, y, dist, fg);
}

uint32_t orc_alpha_blend_u8(uint32_t bg, uint32_t fg) { return alpha_blend_u8(bg, fg); }

void orc_alpha_uv16(const int32_t* e1, const int32_t* e2, const float* coef, const float* remap, const float* uv,
                    float slu, float slv, float* U, float* V, int32_t* mip) {
    FragmentCtx fc = {};
    fc.alpha_test = true;
    for (int k = 0; k < 3; k++) {
        fc.uv[k][0] = uv[k * 2];
        fc.uv[k][1] = uv[k * 2 + 1];
    }
    if (remap) {
        fc.clipped = true;
        for (int k = 0; k < 3; k++) {
            fc.ru[k] = remap[k];
            fc.rv[k] = remap[3 + k];
        }
    }
    float bu[16], bv[16];
    for (int l = 0; l < 16; l++)
        fragment_bary((float)e1[l], (float)e2[l], coef[3], coef[4], coef[5], coef[6], bu[l], bv[l]);
    alpha_uv_tile(fc, slu, slv, bu, bv, U, V, mip);
}

void orc_bary_f64(const float* clip, double su, double sv, double* bary) {
    double px[3], py[3], iw[3];
    for (int k = 0; k < 3; k++) {
        iw[k] = 1.0 / (double)clip[k * 4 + 3];
        px[k] = (double)clip[k * 4 + 0] * iw[k];
        py[k] = (double)clip[k * 4 + 1] * iw[k];
    }
    // screen-space barycentrics, then weight by 1/w and renormalise
    double area = (px[1] - px[0]) * (py[2] - py[0]) - (py[1] - py[0]) * (px[2] - px[0]);
    double l1 = ((su - px[0]) * (py[2] - py[0]) - (sv - py[0]) * (px[2] - px[0])) / area;
    double l2 = ((px[1] - px[0]) * (sv - py[0]) - (py[1] - py[0]) * (su - px[0])) / area;
    double l0 = 1.0 - l1 - l2;
    double q0 = l0 * iw[0], q1 = l1 * iw[1], q2 = l2 * iw[2];
    double s = q0 + q1 + q2;
    bary[0] = q0 / s;
    bary[1] = q1 / s;
    bary[2] = q2 / s;
}

}  // extern "C"

namespace {

// Discrete choices and fp32 values of one pixel of the differentiable
// resolve, in the GPU's operation order.
struct DiffLane {
    bool active;
    uint32_t mi, vi[3];
    float b1, b2;            // barycentrics (constants of the function)
    int32_t tex;             // -1: untextured
    int32_t level;
    bool linear;
    float floorU, floorV;    // wrap period
    int32_t tx, ty;          // nearest texel, or the bilinear cell's corner
    bool cx, cy;             // bilinear: inside the left/top clamp
    bool ex, ey;             // bilinear: right/bottom neighbour exists
};

void diff_tile(const orc_shading_ctx& ctx, const orc_diff_params& p, const uint32_t* ids, const bool* active,
               const float* su, const float* sv, float sw, float sh, DiffLane* L) {
    const orc_meshlets& ms = *ctx.meshlets;
    float gradU[16][2], gradV[16][2], U[16], V[16];
    uint32_t material[16];
    for (int l = 0; l < 16; l++) {
        L[l] = DiffLane{};
        L[l].active = active[l];
        L[l].tex = -1;
        material[l] = UINT32_MAX;
        U[l] = V[l] = 0.0f;
        gradU[l][0] = gradU[l][1] = gradV[l][0] = gradV[l][1] = 0.0f;
        if (!active[l]) continue;
        uint32_t mi = ids[l] / 128, tri = ids[l] % 128;
        L[l].mi = mi;
        float c[3][4], uv[3][2];
        for (int k = 0; k < 3; k++) {
            uint32_t vi = ms.indices[(size_t)mi * 384 + k * 128 + tri];
            L[l].vi[k] = vi;
            const float* P = ms.positions + (size_t)mi * 192;
            transform(p.object_to_clip, P[vi], P[64 + vi], P[128 + vi], c[k]);
            uv[k][0] = p.uv[((size_t)mi * 64 + vi) * 2 + 0];
            uv[k][1] = p.uv[((size_t)mi * 64 + vi) * 2 + 1];
        }
        material[l] = ms.material_id[mi];
        float bary[3], ddx[3], ddy[3];
        intersect_triangle_exact(c, su[l], sv[l], sw, sh, bary, ddx, ddy);
        L[l].b1 = bary[1];
        L[l].b2 = bary[2];
        float t10x = uv[1][0] - uv[0][0], t10y = uv[1][1] - uv[0][1];
        float t20x = uv[2][0] - uv[0][0], t20y = uv[2][1] - uv[0][1];
        U[l] = fmaf(bary[2], t20x, fmaf(bary[1], t10x, uv[0][0]));
        V[l] = fmaf(bary[2], t20y, fmaf(bary[1], t10y, uv[0][1]));
        gradU[l][0] = fmaf(t20x, ddx[2], t10x * ddx[1]);
        gradV[l][0] = fmaf(t20y, ddx[2], t10y * ddx[1]);
        gradU[l][1] = fmaf(t20x, ddy[2], t10x * ddy[1]);
        gradV[l][1] = fmaf(t20y, ddy[2], t10y * ddy[1]);
    }
    // material waterfall with the tile-wide filter vote, as in the resolve
    bool done[16] = {};
    for (int l0 = 0; l0 < 16; l0++) {
        if (!active[l0] || done[l0] || material[l0] == UINT32_MAX) continue;
        const uint32_t mat = material[l0];
        const int32_t ti = ctx.materials[mat].texture;
        if (ti < 0 || (uint32_t)ti >= p.num_textures) {
            for (int l = 0; l < 16; l++) if (active[l] && material[l] == mat) done[l] = true;
            continue;
        }
        const orc_diff_texture& T = p.textures[ti];
        const float su2 = (float)T.width * (float)T.width, sv2 = (float)T.height * (float)T.height;
        int32_t mip[16];
        bool minify = false;
        for (int l = 0; l < 16; l++) {
            float dx = fmaf(gradU[l][0], gradU[l][0], gradV[l][0] * gradV[l][0]) * su2;
            float dy = fmaf(gradU[l][1], gradU[l][1], gradV[l][1] * gradV[l][1]) * sv2;
            float mx = (dx > dy) ? dx : dy;
            mip[l] = wrap_sub((int32_t)fbits(mx), 127 << 23) >> 24;
            if (active[l] && mip[l] > 0) minify = true;
        }
        for (int l = 0; l < 16; l++) {
            if (!active[l] || material[l] != mat) continue;
            done[l] = true;
            DiffLane& D = L[l];
            D.tex = ti;
            D.linear = !minify;
            D.floorU = floorf(U[l]);
            D.floorV = floorf(V[l]);
            const float fu = U[l] - D.floorU, fv = V[l] - D.floorV;
            if (D.linear) {
                D.level = 0;
                const int32_t W = (int32_t)T.width, H = (int32_t)T.height;
                const float xs = fu * (float)W - 0.5f, ys = fv * (float)H - 0.5f;
                D.cx = xs > 0.0f;
                D.cy = ys > 0.0f;
                const float xf = D.cx ? xs : 0.0f, yf = D.cy ? ys : 0.0f;
                D.tx = (int32_t)xf;
                D.ty = (int32_t)yf;
                D.ex = D.tx < W - 1;
                D.ey = D.ty < H - 1;
                if (!D.ex) D.tx = W - 1;
                if (!D.ey) D.ty = H - 1;
            } else {
                int32_t m = mip[l] < 0 ? 0 : mip[l];
                if (m > (int32_t)T.mip_levels - 1) m = (int32_t)T.mip_levels - 1;
                D.level = m;
                const int32_t Wm = (int32_t)(T.width >> m), Hm = (int32_t)(T.height >> m);
                int32_t x = (int32_t)(fu * (float)Wm), y = (int32_t)(fv * (float)Hm);
                D.tx = x < Wm - 1 ? x : Wm - 1;
                D.ty = y < Hm - 1 ? y : Hm - 1;
            }
        }
    }
}

struct DiffEval {
    double w[3];             // d(U,V,col)/d(vertex k)
    double U, V, col[4], val[4];
    uint64_t texel[4];       // bilinear corners 00, 10, 01, 11 (or texel[0] for nearest)
    double tw[4];            // their weights
    int ntexels;
    double dval_du[4], dval_dv[4];
};

void diff_eval(const orc_shading_ctx& ctx, const orc_diff_params& p, const DiffLane& D, uint32_t x, uint32_t y,
               uint32_t W, uint32_t H, DiffEval& E) {
    (void)ctx;
    (void)x;
    (void)y;
    (void)W;
    (void)H;
    const double b[3] = { 1.0 - (double)D.b1 - (double)D.b2, (double)D.b1, (double)D.b2 };
    E.w[0] = 1.0 - b[1] - b[2];
    E.w[1] = b[1];
    E.w[2] = b[2];
    auto uvv = [&](int k, int j) {
        size_t i = ((size_t)D.mi * 64 + D.vi[k]) * 2 + j;
        return p.uv_d ? p.uv_d[i] : (double)p.uv[i];
    };
    auto colv = [&](int k, int j) {
        size_t i = ((size_t)D.mi * 64 + D.vi[k]) * 4 + j;
        return p.color_d ? p.color_d[i] : (double)p.color[i];
    };
    auto tex = [&](uint64_t t, int j) { return p.texels_d ? p.texels_d[t * 4 + j] : (double)p.texels[t * 4 + j]; };
    E.U = uvv(0, 0) + b[1] * (uvv(1, 0) - uvv(0, 0)) + b[2] * (uvv(2, 0) - uvv(0, 0));
    E.V = uvv(0, 1) + b[1] * (uvv(1, 1) - uvv(0, 1)) + b[2] * (uvv(2, 1) - uvv(0, 1));
    for (int j = 0; j < 4; j++) E.col[j] = colv(0, j) + b[1] * (colv(1, j) - colv(0, j)) + b[2] * (colv(2, j) - colv(0, j));
    E.ntexels = 0;
    for (int j = 0; j < 4; j++) E.dval_du[j] = E.dval_dv[j] = 0.0;
    if (D.tex < 0) {
        for (int j = 0; j < 4; j++) E.val[j] = 1.0;
        return;
    }
    const orc_diff_texture& T = p.textures[D.tex];
    const uint64_t base = T.level_offset[D.level];
    if (!D.linear) {
        const uint64_t Wm = T.width >> D.level;
        E.ntexels = 1;
        E.texel[0] = base + (uint64_t)D.ty * Wm + (uint64_t)D.tx;
        E.tw[0] = 1.0;
        for (int j = 0; j < 4; j++) E.val[j] = tex(E.texel[0], j);
        return;
    }
    const double fu = E.U - (double)D.floorU, fv = E.V - (double)D.floorV;
    const double xf = D.cx ? fu * T.width - 0.5 : 0.0, yf = D.cy ? fv * T.height - 0.5 : 0.0;
    const double fx = D.ex ? xf - D.tx : 0.0, fy = D.ey ? yf - D.ty : 0.0;
    const uint64_t x0 = (uint64_t)D.tx, y0 = (uint64_t)D.ty;
    const uint64_t x1 = D.ex ? x0 + 1 : x0, y1 = D.ey ? y0 + 1 : y0;
    E.ntexels = 4;
    E.texel[0] = base + y0 * T.width + x0;
    E.texel[1] = base + y0 * T.width + x1;
    E.texel[2] = base + y1 * T.width + x0;
    E.texel[3] = base + y1 * T.width + x1;
    E.tw[0] = (1.0 - fx) * (1.0 - fy);
    E.tw[1] = fx * (1.0 - fy);
    E.tw[2] = (1.0 - fx) * fy;
    E.tw[3] = fx * fy;
    const double sx = (D.cx && D.ex) ? (double)T.width : 0.0, sy = (D.cy && D.ey) ? (double)T.height : 0.0;
    for (int j = 0; j < 4; j++) {
        double t00 = tex(E.texel[0], j), t10 = tex(E.texel[1], j), t01 = tex(E.texel[2], j), t11 = tex(E.texel[3], j);
        double l0 = t00 + fx * (t10 - t00), l1 = t01 + fx * (t11 - t01);
        E.val[j] = l0 + fy * (l1 - l0);
        E.dval_du[j] = sx * ((1.0 - fy) * (t10 - t00) + fy * (t11 - t01));
        E.dval_dv[j] = sy * (l1 - l0);
    }
}

template <typename F>
void diff_for_each(const orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_diff_params* p, F&& fn) {
    const uint32_t W = fb->width, H = fb->height;
    const float sw = (float)W, sh = (float)H;
    const float scaleX = 2.0f / sw, scaleY = 2.0f / sh;
    const float centerX = fmaf(0.5f, scaleX, -1.0f), centerY = fmaf(0.5f, scaleY, -1.0f);
    const float* depth = (const float*)(fb->data + fb->layer_stride);
    for (uint32_t y0 = 0; y0 < H; y0 += 4) {
        for (uint32_t x0 = 0; x0 < W; x0 += 4) {
            const uint32_t off = pixel_offset(W, x0, y0);
            bool active[16], any = false;
            float sx[16], sy[16];
            uint32_t ids[16];
            for (int l = 0; l < 16; l++) {
                active[l] = !(depth[off + l] <= 0.0f);
                any |= active[l];
                ids[l] = fb->data[off + l];
                sx[l] = fmaf((float)(int32_t)(x0 + (l & 3)), scaleX, centerX);
                sy[l] = fmaf((float)(int32_t)(y0 + (l >> 2)), scaleY, centerY);
            }
            if (!any) continue;
            DiffLane L[16];
            diff_tile(*ctx, *p, ids, active, sx, sy, sw, sh, L);
            for (int l = 0; l < 16; l++) {
                if (!L[l].active) continue;
                const uint32_t x = x0 + (l & 3), y = y0 + (l >> 2);
                DiffEval E;
                diff_eval(*ctx, *p, L[l], x, y, W, H, E);
                fn(L[l], E, (size_t)y * W + x);
            }
        }
    }
}

}  // namespace

extern "C" {

void orc_diff_forward(const orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_diff_params* p, double* out) {
    memset(out, 0, sizeof(double) * 4 * fb->width * fb->height);
    diff_for_each(fb, ctx, p, [&](const DiffLane&, const DiffEval& E, size_t pix) {
        for (int j = 0; j < 4; j++) out[pix * 4 + j] = E.val[j] * E.col[j];
    });
}

void orc_diff_backward(const orc_framebuffer* fb, const orc_shading_ctx* ctx, const orc_diff_params* p,
                       const double* grad_out, double* grad_texels, double* grad_uv, double* grad_color) {
    diff_for_each(fb, ctx, p, [&](const DiffLane& D, const DiffEval& E, size_t pix) {
        const double* g = grad_out + pix * 4;
        double gU = 0.0, gV = 0.0;
        for (int j = 0; j < 4; j++) {
            const double gval = g[j] * E.col[j];
            for (int t = 0; t < E.ntexels; t++) grad_texels[E.texel[t] * 4 + j] += E.tw[t] * gval;
            gU += gval * E.dval_du[j];
            gV += gval * E.dval_dv[j];
        }
        for (int k = 0; k < 3; k++) {
            const size_t v = (size_t)D.mi * 64 + D.vi[k];
            grad_uv[v * 2 + 0] += E.w[k] * gU;
            grad_uv[v * 2 + 1] += E.w[k] * gV;
            for (int j = 0; j < 4; j++) grad_color[v * 4 + j] += E.w[k] * g[j] * E.val[j];
        }
    });
}

}  // extern "C"
