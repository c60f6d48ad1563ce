// Toolchain probe: how clang lowers the reference's floating-point expression
// shapes under the reference's flags (-O2 -march=native
// -mprefer-vector-width=512 -ffast-math, src/SwRast/CMakeLists.txt:6).
//
// Each function below evaluates one expression shape of the reference on
// 16-lane clang ext-vectors, the way the reference writes it (its own 16-lane
// vector types, SIMD.h:16-19). The code is written here from the formulas; it
// is not the reference's source. Compiled with those flags, the functions
// show which divisions become VRCP14PS + Newton step and which products are
// contracted into FMAs. The oracle states the resulting operation order
// explicitly; tests/test_oracle_fastmath.py runs these probes on random inputs
// and requires the oracle to match them bit for bit.
//
// Only tests load this library; it is test infrastructure, never product code.
#include <cstdint>
#include <cstring>
#include <immintrin.h>

typedef float vf __attribute__((ext_vector_type(16)));
typedef int32_t vi __attribute__((ext_vector_type(16)));

static inline vf ld(const float* p) { vf v; memcpy(&v, p, 64); return v; }
static inline vi ldi(const int32_t* p) { vi v; memcpy(&v, p, 64); return v; }
static inline void st(float* p, vf v) { memcpy(p, &v, 64); }
static inline void sti(int32_t* p, vi v) { memcpy(p, &v, 64); }
static inline vf vfma(vf a, vf b, vf c) { return __builtin_elementwise_fma(a, b, c); }
static inline vi rne(vf x) { return (vi)_mm512_cvtps_epi32((__m512)x); }

extern "C" {

// 1/w (SIMD.h:474)
void probe_recip(const float* w, float* out, int n) {
    for (int i = 0; i < n; i += 16) st(out + i, 1.0f / ld(w + i));
}
// generic a/b on vectors
void probe_div(const float* a, const float* b, float* out, int n) {
    for (int i = 0; i < n; i += 16) st(out + i, ld(a + i) / ld(b + i));
}
// VRCP14PS / VRSQRT14PS (SIMD.h:294-295)
void probe_rcp14(const float* x, float* out, int n) {
    for (int i = 0; i < n; i += 16) st(out + i, (vf)_mm512_rcp14_ps((__m512)ld(x + i)));
}
void probe_rsqrt14(const float* x, float* out, int n) {
    for (int i = 0; i < n; i += 16) st(out + i, (vf)_mm512_rsqrt14_ps((__m512)ld(x + i)));
}

// Triangle setup shape (Rasterizer.cpp:256-289 with SIMD.h:473-476):
// perspective divide of three clip-space vertices, signed area with the
// cull-mode flip, and RNE snapping to 1/16 px.
// in: clip[3][4][n] (vertex, xyzw, lane); cullMode 0=None 1=FrontCCW 2=FrontCW
// out: X[3][n], Y[3][n] snapped ints, Z[3][n], W[3][n] floats, det[n]
void probe_setup(const float* clip, int n, float fsx, float fsy, int cullMode,
                 int32_t* X, int32_t* Y, float* Z, float* W, float* det) {
    for (int i = 0; i < n; i += 16) {
        vf x[3], y[3], z[3], rw[3];
        for (int k = 0; k < 3; k++) {
            vf cx = ld(clip + (k * 4 + 0) * n + i), cy = ld(clip + (k * 4 + 1) * n + i);
            vf cz = ld(clip + (k * 4 + 2) * n + i), cw = ld(clip + (k * 4 + 3) * n + i);
            rw[k] = 1.0f / cw;
            x[k] = cx * rw[k]; y[k] = cy * rw[k]; z[k] = cz * rw[k];
        }
        vf d = (x[2] - x[0]) * (y[1] - y[0]) - (x[0] - x[1]) * (y[0] - y[2]);
        if (cullMode != 1) {
            vi flip = (cullMode == 2) ? vi(-1) : (d < 0);
            d = flip ? -d : d;
        }
        st(det + i, d);
        for (int k = 0; k < 3; k++) {
            sti(X + k * n + i, rne(x[k] * fsx));
            sti(Y + k * n + i, rne(y[k] * fsy));
            st(Z + k * n + i, z[k]);
            st(W + k * n + i, rw[k]);
        }
    }
}

// Edge setup float part (Rasterizer.cpp:320-328): rcpArea from the integer
// doubled area, depth plane and scaled 1/w terms.
void probe_edge(const int32_t* det, const float* Z, const float* W, int n,
                float* Z10, float* Z20, float* W0S, float* W1S, float* W2S) {
    for (int i = 0; i < n; i += 16) {
        vf rcpArea = 16.0f / __builtin_convertvector(ldi(det + i), vf);
        vf z0 = ld(Z + i), z1 = ld(Z + n + i), z2 = ld(Z + 2 * n + i);
        st(Z10 + i, (z1 - z0) * rcpArea);
        st(Z20 + i, (z2 - z0) * rcpArea);
        st(W0S + i, ld(W + i) * rcpArea);
        st(W1S + i, ld(W + n + i) * rcpArea);
        st(W2S + i, ld(W + 2 * n + i) * rcpArea);
    }
}

// Per-fragment depth and perspective-correct barycentrics
// (Rasterizer.h:293-319, unclipped case).
void probe_fragment(const int32_t* e1, const int32_t* e2, const float* coef, int n,
                    float* depth, float* bu, float* bv) {
    // coef: Z0, Z10, Z20, W0, W0S, W1S, W2S (scalars broadcast)
    vf Z0 = coef[0], Z10 = coef[1], Z20 = coef[2], W0 = coef[3], W0S = coef[4], W1S = coef[5], W2S = coef[6];
    for (int i = 0; i < n; i += 16) {
        vf u = __builtin_convertvector(ldi(e1 + i), vf);
        vf v = __builtin_convertvector(ldi(e2 + i), vf);
        st(depth + i, vfma(u, Z10, vfma(v, Z20, Z0)));
        vf pw0 = vfma(u + v, -W0S, W0);
        vf w = vfma(u, W1S, vfma(v, W2S, pw0));
        vf r = (vf)_mm512_rcp14_ps((__m512)w);
        r *= vfma(-w, r, 2.0f);
        u *= W1S * r;
        v *= W2S * r;
        st(bu + i, u);
        st(bv + i, v);
    }
}

// Alpha-test UVs of one 4x4 tile (Shading.cpp:316-327 with
// Rasterizer.h:293-319 and Texture.h:260-280, 403-410): perspective-correct
// barycentrics, the clipped-triangle remap, BaryLerp of the vertex UVs, the
// lerp-scaled UVs and their tile derivatives, and the implicit mip level.
// coef: W0, W0S, W1S, W2S; remap: ClippedU[3], ClippedV[3] or null; uv[3][2].
void probe_alpha_uv(const int32_t* e1, const int32_t* e2, const float* coef, const float* remap, const float* uv,
                    float slu, float slv, float* U, float* V, int32_t* mip) {
    vf W0 = coef[0], W0S = coef[1], W1S = coef[2], W2S = coef[3];
    vf u = __builtin_convertvector(ldi(e1), vf);
    vf v = __builtin_convertvector(ldi(e2), vf);
    vf pw0 = vfma(u + v, -W0S, W0);
    vf w = vfma(u, W1S, vfma(v, W2S, pw0));
    vf r = (vf)_mm512_rcp14_ps((__m512)w);
    r *= vfma(-w, r, 2.0f);
    u *= W1S * r;
    v *= W2S * r;
    if (remap) {
        vf cu = vfma(u, (vf)remap[1], vfma(v, (vf)remap[2], (vf)remap[0]));
        vf cv = vfma(u, (vf)remap[4], vfma(v, (vf)remap[5], (vf)remap[3]));
        u = cu, v = cv;
    }
    vf b0 = 1 - u - v;
    vf Ux = vfma((vf)uv[0], b0, vfma((vf)uv[2], u, (vf)uv[4] * v));
    vf Vy = vfma((vf)uv[1], b0, vfma((vf)uv[3], u, (vf)uv[5] * v));
    st(U, Ux);
    st(V, Vy);
    __m512 su = (__m512)(Ux * slu), sv = (__m512)(Vy * slv);
    auto ddx = [](__m512 p) { return _mm512_sub_ps(_mm512_shuffle_ps(p, p, 0b11110101), _mm512_shuffle_ps(p, p, 0b10100000)); };
    auto ddy = [](__m512 p) {
        return _mm512_sub_ps(_mm512_shuffle_f32x4(p, p, 0b11110101), _mm512_shuffle_f32x4(p, p, 0b10100000));
    };
    vf gx = (vf)ddx(su), gy = (vf)ddx(sv), gz = (vf)ddy(su), gw = (vf)ddy(sv);
    vf dx = vfma(gx, gx, gy * gy), dy = vfma(gz, gz, gw * gw);
    vf mx = (vf)_mm512_max_ps((__m512)dx, (__m512)dy);
    vi m = ((((vi)mx) - (127 << 23)) >> 23) >> 1;
    sti(mip, m - 8);
}

// Normal mapping of the resolve (Shading.cpp:556-572): bitangent = cross(N, T)
// (SIMD.h:431-437) with the handedness sign, tangent-space normal from the
// two 8-bit channels of the NMR texel with z = approx_sqrt(1 - x^2 - y^2)
// (SIMD.h:296), then normalize(x*T + y*B + z*N) (SIMD.h:439).
// in: n[3][16], t[3][16] (world-space normal and tangent), hand[16] (0 or
// 1<<31), nmr[16]; out[3][16]
void probe_normal_map(const float* n, const float* t, const int32_t* hand, const int32_t* nmr, float* out) {
    vf nx = ld(n), ny = ld(n + 16), nz = ld(n + 32);
    vf tx = ld(t), ty = ld(t + 16), tz = ld(t + 32);
    vf bx = vfma(ny, tz, -nz * ty);
    vf by = vfma(nz, tx, -nx * tz);
    vf bz = vfma(nx, ty, -ny * tx);
    vi h = ldi(hand);
    bx = (vf)((vi)bx ^ h);
    by = (vf)((vi)by ^ h);
    bz = (vf)((vi)bz ^ h);
    vi p = ldi(nmr);
    vf sx = __builtin_convertvector(p & 255, vf) * (1.0f / 127.5f) - 1.0f;
    vf sy = __builtin_convertvector((p >> 8) & 255, vf) * (1.0f / 127.5f) - 1.0f;
    vf s = 1.0f - (sx * sx + sy * sy);
    vf sz = (vf)_mm512_rsqrt14_ps((__m512)s) * s;
    vf rx = sx * tx + sy * bx + sz * nx;
    vf ry = sx * ty + sy * by + sz * ny;
    vf rz = sx * tz + sy * bz + sz * nz;
    vf r = (vf)_mm512_rsqrt14_ps((__m512)vfma(rx, rx, vfma(ry, ry, rz * rz)));
    st(out, rx * r);
    st(out + 16, ry * r);
    st(out + 32, rz * r);
}

// Light sprite placement of the resolve (Shading.cpp:692-706), scalar float
// math as glm 1.0.1 writes it: mat4 * vec4 = (m0*v.x + m1*v.y) + (m2*v.z +
// m3*v.w) per component, then the clip test, depth, screen position, radius
// and the tile range. m: column-major 4x4; out: depth, centre x, centre y,
// radius; tiles: start x, start y, end x, end y; returns 0 when culled.
int probe_sprite_host(const float* m, const float* pos, int W, int H, float* out, int32_t* tiles) {
    float c[4];
    for (int r = 0; r < 4; r++) {
        float mul0 = m[0 * 4 + r] * pos[0], mul1 = m[1 * 4 + r] * pos[1];
        float mul2 = m[2 * 4 + r] * pos[2], mul3 = m[3 * 4 + r] * 1.0f;
        float add0 = mul0 + mul1, add1 = mul2 + mul3;
        c[r] = add0 + add1;
    }
    float mx = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(c[0]), __builtin_fabsf(c[1])), __builtin_fabsf(c[2]));
    if (mx > c[3]) return 0;
    float depth = c[2] / c[3];
    float sx = c[0] / c[3], sy = c[1] / c[3];
    sx = sx * 0.5f + 0.5f;
    sy = sy * 0.5f + 0.5f;
    float radius = ((float)(W > H ? W : H) / 30.0f) / c[3];
    float cx = sx * (float)W, cy = sy * (float)H;
    int x0 = (int)(cx - radius), y0 = (int)(cy - radius);
    int x1 = (int)(cx + radius), y1 = (int)(cy + radius);
    tiles[0] = (x0 > 0 ? x0 : 0) & ~3;
    tiles[1] = (y0 > 0 ? y0 : 0) & ~3;
    tiles[2] = x1 < W ? x1 : W;
    tiles[3] = y1 < H ? y1 : H;
    out[0] = depth; out[1] = cx; out[2] = cy; out[3] = radius;
    return 1;
}

// Light sprite colour of one 4x4 tile (Shading.cpp:712-724): distance to the
// centre, alpha 1 - a^2 with a = 1 - (-d / r^2), packed with the light colour
// (RGBA8u::Pack, Texture.h:55-67). out: dist[16] floats, fg[16] packed.
void probe_sprite_tile(int x, int y, float cx, float cy, float radius, const float* color, float* dist,
                       int32_t* fg) {
    vi lane = { 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15 };
    vi tx = x + (lane & 3), ty = y + (lane >> 2);
    vf rx = __builtin_convertvector(tx, vf) + 0.5f - cx;
    vf ry = __builtin_convertvector(ty, vf) + 0.5f - cy;
    vf d = vfma(rx, rx, ry * ry) - radius * radius;
    vf a = 1 - (-d / (radius * radius));
    vf alpha = 1 - a * a;
    vi r = rne((vf)color[0] * 255.0f), g = rne((vf)color[1] * 255.0f), b = rne((vf)color[2] * 255.0f);
    vi al = rne(alpha * 255.0f);
    auto sat = [](vi v) {
        v = v < -32768 ? -32768 : v;
        v = v > 32767 ? 32767 : v;
        v = v < 0 ? 0 : v;
        return v > 255 ? 255 : v;
    };
    st(dist, d);
    sti(fg, sat(r) | (sat(g) << 8) | (sat(b) << 16) | (sat(al) << 24));
}

// Occlusion part of CullMeshlets (Shading.cpp:811-832) with ProjectSphere
// (Shading.cpp:264-279) for 16 bound spheres: view-space centre (simd::mul,
// SIMD.h:457-464), projected box, mip level, texel range and sphere depth.
// in: sphere[4][16] (x, y, z, r), m column-major object-to-previous-view,
// scale, proj (znear, f/ar, -f), fs (frame size), mipLevels;
// out: mask[16], bb[4][16], mip[16], rect[4][16] (x0, y0, x1, y1), depth[16]
void probe_hiz_sphere(const float* sphere, const float* m, float scale, const float* proj, const float* fs,
                      int mipLevels, int32_t* mask, float* bb, int32_t* mip, int32_t* rect, float* depth) {
    vf px = ld(sphere), py = ld(sphere + 16), pz = ld(sphere + 32), pr = ld(sphere + 48);
    vf one = 1.0f;
    vf cx = vfma(px, m[0], vfma(py, m[4], vfma(pz, m[8], one * m[12])));
    vf cy = vfma(px, m[1], vfma(py, m[5], vfma(pz, m[9], one * m[13])));
    vf cz = vfma(px, m[2], vfma(py, m[6], vfma(pz, m[10], one * m[14])));
    vf r = pr * scale;
    // ProjectSphere
    vf c_z = cz * -1.0f;
    vi msk = c_z >= r + proj[0];
    vf crx = cx * r, cry = cy * r, crz = c_z * r;
    vf vx = (vf)_mm512_rsqrt14_ps((__m512)(cx * cx + c_z * c_z - r * r)) * (cx * cx + c_z * c_z - r * r);
    vf vy = (vf)_mm512_rsqrt14_ps((__m512)(cy * cy + c_z * c_z - r * r)) * (cy * cy + c_z * c_z - r * r);
    vf rcp0 = (vf)_mm512_rcp14_ps((__m512)(vx * c_z + crx));
    vf rcp1 = (vf)_mm512_rcp14_ps((__m512)(vy * c_z - cry));
    vf rcp2 = (vf)_mm512_rcp14_ps((__m512)(vx * c_z - crx));
    vf rcp3 = (vf)_mm512_rcp14_ps((__m512)(vy * c_z + cry));
    vf b0 = (vx * cx - crz) * rcp0 * (proj[1] * 0.5f) + 0.5f;
    vf b1 = (vy * cy + crz) * rcp1 * (proj[2] * 0.5f) + 0.5f;
    vf b2 = (vx * cx + crz) * rcp2 * (proj[1] * 0.5f) + 0.5f;
    vf b3 = (vy * cy - crz) * rcp3 * (proj[2] * 0.5f) + 0.5f;
    vf sx = (b2 - b0) * fs[0], sy = (b3 - b1) * fs[1];
    vf mx = (vf)_mm512_max_ps((__m512)sx, (__m512)sy);
    vi lv = (((vi)mx) - (127 << 23)) >> 23;
    lv = lv < 1 ? 1 : lv;
    lv = lv > mipLevels - 1 ? mipLevels - 1 : lv;
    vi x0 = __builtin_convertvector(b0 * fs[0], vi);
    vi y0 = __builtin_convertvector(b1 * fs[1], vi);
    vi x1 = __builtin_convertvector(b2 * fs[0], vi);
    vi y1 = __builtin_convertvector(b3 * fs[1], vi);
    x0 = (x0 < 0 ? 0 : x0) >> lv;
    y0 = (y0 < 0 ? 0 : y0) >> lv;
    x1 = (x1 > (int)fs[0] - 1 ? (int)fs[0] - 1 : x1) >> lv;
    y1 = (y1 > (int)fs[1] - 1 ? (int)fs[1] - 1 : y1) >> lv;
    vf d = proj[0] / (-cz - r);
    sti(mask, msk);
    st(bb, b0); st(bb + 16, b1); st(bb + 32, b2); st(bb + 48, b3);
    sti(mip, lv);
    sti(rect, x0); sti(rect + 16, y0); sti(rect + 32, x1); sti(rect + 48, y1);
    st(depth, d);
}

// Screen UV of a 4x4 tile in a full-screen pass (Rasterizer.h:226-237).
void probe_tile_uv(const int32_t* px, int n, float size, float* out) {
    float scale = 2.0f / size;
    float center = 0.5f * scale - 1.0f;
    for (int i = 0; i < n; i += 16) {
        vf p = __builtin_convertvector(ldi(px + i), vf);
        st(out + i, p * scale + center);
    }
}

// Analytic barycentrics and screen derivatives of a triangle given clip
// positions and a screen UV (filmic-worlds method, Shading.cpp:417-464).
// in: v[3][4][n], suv[2][n]; out: bary[3][n], ddx[3][n], ddy[3][n]
void probe_intersect(const float* v, const float* suv, int n, float sw, float sh,
                     float* bary, float* ddxo, float* ddyo) {
    for (int i = 0; i < n; i += 16) {
        vf x0 = ld(v + 0 * n + i), y0 = ld(v + 1 * n + i), w0 = ld(v + 3 * n + i);
        vf x1 = ld(v + 4 * n + i), y1 = ld(v + 5 * n + i), w1 = ld(v + 7 * n + i);
        vf x2 = ld(v + 8 * n + i), y2 = ld(v + 9 * n + i), w2 = ld(v + 11 * n + i);
        vf su = ld(suv + i), sv = ld(suv + n + i);
        vf iw0 = 1.0f / w0, iw1 = 1.0f / w1, iw2 = 1.0f / w2;
        vf p0x = x0 * iw0, p0y = y0 * iw0;
        vf p1x = x1 * iw1, p1y = y1 * iw1;
        vf p2x = x2 * iw2, p2y = y2 * iw2;
        vf m0x = p2x - p1x, m0y = p2y - p1y, m1x = p0x - p1x, m1y = p0y - p1y;
        vf invDet = 1.0f / (m0x * m1y - m1x * m0y);
        vf dx0 = (p1y - p2y) * (invDet * iw0), dx1 = (p2y - p0y) * (invDet * iw1), dx2 = (p0y - p1y) * (invDet * iw2);
        vf dy0 = (p2x - p1x) * (invDet * iw0), dy1 = (p0x - p2x) * (invDet * iw1), dy2 = (p1x - p0x) * (invDet * iw2);
        vf dxs = dx0 + dx1 + dx2;
        vf dys = dy0 + dy1 + dy2;
        vf r0x = su - p0x, r0y = sv - p0y;
        vf interpInvW = iw0 + r0x * dxs + r0y * dys;
        vf interpW = 1.0f / interpInvW;
        vf by = interpW * (r0x * dx1 + r0y * dy1);
        vf bz = interpW * (r0x * dx2 + r0y * dy2);
        vf bx = 1 - by - bz;
        float kx = 2.0f / sw, ky = -(2.0f / sh);
        dx0 *= kx; dx1 *= kx; dx2 *= kx;
        dy0 *= ky; dy1 *= ky; dy2 *= ky;
        dxs *= kx; dys *= ky;
        vf wdx = 1.0f / (interpInvW + dxs);
        vf wdy = 1.0f / (interpInvW + dys);
        vf ox0 = wdx * (bx * interpInvW + dx0) - bx, ox1 = wdx * (by * interpInvW + dx1) - by, ox2 = wdx * (bz * interpInvW + dx2) - bz;
        vf oy0 = wdy * (bx * interpInvW + dy0) - bx, oy1 = wdy * (by * interpInvW + dy1) - by, oy2 = wdy * (bz * interpInvW + dy2) - bz;
        st(bary + i, bx); st(bary + n + i, by); st(bary + 2 * n + i, bz);
        st(ddxo + i, ox0); st(ddxo + n + i, ox1); st(ddxo + 2 * n + i, ox2);
        st(ddyo + i, oy0); st(ddyo + n + i, oy1); st(ddyo + 2 * n + i, oy2);
    }
}

// Texture coordinate interpolation, UV gradients and mip level
// (Shading.cpp:516-539, Texture.h:276-280 and SIMD.h:426).
// in: t[3][2][n] (vertex uv), bary[3][n], ddx[3][n], ddy[3][n]
// out: uv[2][n], mip[n]
void probe_uvgrad(const float* t, const float* bary, const float* ddx, const float* ddy, int n,
                  float su, float sv, float* uv, int32_t* mip) {
    for (int i = 0; i < n; i += 16) {
        vf t0x = ld(t + 0 * n + i), t0y = ld(t + 1 * n + i);
        vf t10x = ld(t + 2 * n + i) - t0x, t10y = ld(t + 3 * n + i) - t0y;
        vf t20x = ld(t + 4 * n + i) - t0x, t20y = ld(t + 5 * n + i) - t0y;
        vf by = ld(bary + n + i), bz = ld(bary + 2 * n + i);
        vf dxy = ld(ddx + n + i), dxz = ld(ddx + 2 * n + i);
        vf dyy = ld(ddy + n + i), dyz = ld(ddy + 2 * n + i);
        vf u = t0x + t10x * by + t20x * bz;
        vf v = t0y + t10y * by + t20y * bz;
        vf gx = t10x * dxy + t20x * dxz, gy = t10y * dxy + t20y * dxz;
        vf gz = t10x * dyy + t20x * dyz, gw = t10y * dyy + t20y * dyz;
        vf dx = vfma(gx, gx, gy * gy) * (su * su);
        vf dy = vfma(gz, gz, gw * gw) * (sv * sv);
        vf mx = __builtin_elementwise_maxnum(dx, dy);
        vi l = (((vi)mx) - (127 << 23)) >> 23;
        st(uv + i, u); st(uv + n + i, v);
        sti(mip + i, l >> 1);
    }
}

}  // extern "C"

// Overdraw debug colours (ResolveDebug, Shading.cpp:759-766, with
// ColormapTurbo, Shading.cpp:249-260: the published polynomial fit of the
// Turbo colormap, Horner form). in: counter words (pixel count << 16 |
// quad-helper count); quad = 0: pixels / 30, else (pixels + helpers * 0.5) / 30.
// out: rgb[3][n].
static const float kTurbo[18] = {   // double literals rounded to float, as a float array of them is
    0.13572138, 4.61539260,  -42.66032258, 132.13108234, -152.94239396, 59.28637943,
    0.09140261, 2.19418839,  4.84296658,   -14.18503333, 4.27729857,    2.82956604,
    0.10667330, 12.64194608, -60.58204836, 110.36276771, -89.90310912,  27.34824973,
};
extern "C" void probe_overdraw_colour(const uint32_t* data, int quad, float* out, int n) {
    typedef uint32_t vu __attribute__((ext_vector_type(16)));
    for (int i = 0; i < n; i += 16) {
        vu d;
        memcpy(&d, data + i, 64);
        vf pix = __builtin_convertvector(d >> 16, vf);
        vf x;
        if (quad) {
            vf frag = __builtin_convertvector(d & 0xFFFFu, vf);
            x = (pix + frag * 0.5f) / 30.0f;
        } else {
            x = pix / 30.0f;
        }
        for (int k = 0; k < 3; k++) {
            const float* c = kTurbo + 6 * k;
            st(out + k * n + i, x * (x * (x * (x * (x * c[5] + c[4]) + c[3]) + c[2]) + c[1]) + c[0]);
        }
    }
}

// G-buffer packing of FS_EncodeGBuffer (Shading.cpp:397-401): MapOctahedron
// (Texture.h:282-288) of the remapped normal, two 10-bit channels
// conv<uint32_t>(oct * 1023 + 0.5) (truncating), and the NMR texel's
// metallic / roughness bytes cut to 6 bits. n: [3][16]; nmr: [16]; out: [16].
extern "C" void probe_gbuffer_pack(const float* n, const int32_t* nmr, int32_t* out) {
    typedef uint32_t vu __attribute__((ext_vector_type(16)));
    vf x = ld(n), y = ld(n + 16), z = ld(n + 32);
    vf w = (vf)_mm512_rcp14_ps((__m512)(__builtin_elementwise_abs(x) + __builtin_elementwise_abs(y) +
                                        __builtin_elementwise_abs(z)));
    vf t = __builtin_elementwise_maxnum(-z * w, (vf)0.0f);
    vf sx = (vf)((vu)t ^ ((vu)x & 0x80000000u));
    vf sy = (vf)((vu)t ^ ((vu)y & 0x80000000u));
    vf u = vfma(x, w, sx);
    vf v = vfma(y, w, sy);
    u = u * 0.5f + 0.5f;
    v = v * 0.5f + 0.5f;
    vu p;
    memcpy(&p, nmr, 64);
    vu r = __builtin_convertvector(u * 1023.0f + 0.5f, vu);
    r |= __builtin_convertvector(v * 1023.0f + 0.5f, vu) << 10;
    r |= (p >> 18 & 0x3F) << 20;
    r |= (p >> 26 & 0x3F) << 26;
    memcpy(out, &r, 64);
}

// Skybox coordinates of Resolve (Shading.cpp:675-678): SampleOctLevel
// (Texture.h:467-480) maps the view direction with MapOctahedron
// (Texture.h:282-288) and SampleLevel scales and rounds (Texture.h:412-416).
// dir: [3][16]; slu, slv: ScaleLerpU/V (run-time values); out ix, iy: [16].
extern "C" void probe_sky_coords(const float* dir, float slu, float slv, int32_t* ix, int32_t* iy) {
    typedef uint32_t vu __attribute__((ext_vector_type(16)));
    vf x = ld(dir), y = ld(dir + 16), z = ld(dir + 32);
    vf w = (vf)_mm512_rcp14_ps((__m512)(__builtin_elementwise_abs(x) + __builtin_elementwise_abs(y) +
                                        __builtin_elementwise_abs(z)));
    vf t = __builtin_elementwise_maxnum(-z * w, (vf)0.0f);
    vf u = vfma(x, w, (vf)((vu)t ^ ((vu)x & 0x80000000u)));
    vf v = vfma(y, w, (vf)((vu)t ^ ((vu)y & 0x80000000u)));
    u = u * 0.5f + 0.5f;
    v = v * 0.5f + 0.5f;
    sti(ix, rne(u * slu));
    sti(iy, rne(v * slv));
}

// Float bilinear filter of SampleLinear for R11G11B10f texels
// (Texture.h:504-573, 135-185): 8-bit fractions scaled by 1/256, the x weight
// zeroed past the last column, F11/F11/F10 unpacked by bit moves, two row
// lerps and one column lerp. d: [4][16] texels 00, 10, 01, 11; ixf, iyf:
// fixed-point coordinates; inx: x inbound mask; out: rgb [3][16].
extern "C" void probe_env_bilinear(const uint32_t* d, const int32_t* ixf, const int32_t* iyf, const int32_t* inx,
                                   float* out) {
    typedef uint32_t vu __attribute__((ext_vector_type(16)));
    vu t[4];
    for (int k = 0; k < 4; k++) memcpy(&t[k], d + 16 * k, 64);
    vi fxi = ldi(ixf), fyi = ldi(iyf), in = ldi(inx);
    const float fracScale = 1.0f / 256;
    vf fx = __builtin_convertvector(fxi & 255, vf) * fracScale;
    vf fy = __builtin_convertvector(fyi & 255, vf) * fracScale;
    fx = in ? fx : (vf)0.0f;
    for (int c = 0; c < 3; c++) {
        vf col[4];
        for (int k = 0; k < 4; k++) {
            vu x = c == 0 ? t[k] >> 21 : c == 1 ? t[k] >> 10 : t[k];
            vu b = c == 2 ? ((x << 18) & 0x0FFC0000u) + 0x38000000u : ((x << 17) & 0x0FFE0000u) + 0x38000000u;
            col[k] = (vf)b;
        }
        vf rowA = col[0] + (col[1] - col[0]) * fx;
        vf rowB = col[2] + (col[3] - col[2]) * fx;
        st(out + 16 * c, rowA + (rowB - rowA) * fy);
    }
}

// GenerateMip's box filter (Texture.h:577-596): four texels unpacked, averaged
// as (c00 + c10 + c01 + c11) * 0.25f on four-component vectors, packed back.
typedef uint32_t vu __attribute__((ext_vector_type(16)));
struct vf4 { vf x, y, z, w; };
static inline vu ldu(const uint32_t* p) { vu v; memcpy(&v, p, 64); return v; }
static inline void stu(uint32_t* p, vu v) { memcpy(p, &v, 64); }
static inline vf4 add4(vf4 a, vf4 b) { return { a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w }; }
static inline vf4 mul4(vf4 a, float s) { return { a.x * s, a.y * s, a.z * s, a.w * s }; }

// RGBA8u::Unpack (Texture.h:28-36)
static inline vf4 unpack_rgba8(vu p) {
    const float scale = 1.0f / 255;
    return { __builtin_convertvector((p >> 0) & 255, vf) * scale, __builtin_convertvector((p >> 8) & 255, vf) * scale,
             __builtin_convertvector((p >> 16) & 255, vf) * scale, __builtin_convertvector((p >> 24) & 255, vf) * scale };
}
// RGBA8u::Pack (Texture.h:55-67): RNE conversion, then the signed 16-bit and
// unsigned 8-bit saturations of the two packs
static inline vu pack_chan(vf v) {
    vi i = rne(v * 255.0f);
    i = i < -32768 ? -32768 : i;
    i = i > 32767 ? 32767 : i;
    i = i < 0 ? 0 : i;
    i = i > 255 ? 255 : i;
    return (vu)i;
}

extern "C" void probe_mip_rgba8(const uint32_t* c00, const uint32_t* c10, const uint32_t* c01, const uint32_t* c11,
                                uint32_t* out, int n) {
    for (int i = 0; i < n; i += 16) {
        vf4 avg = mul4(add4(add4(add4(unpack_rgba8(ldu(c00 + i)), unpack_rgba8(ldu(c10 + i))),
                                 unpack_rgba8(ldu(c01 + i))), unpack_rgba8(ldu(c11 + i))), 0.25f);
        stu(out + i, pack_chan(avg.x) | (pack_chan(avg.y) << 8) | (pack_chan(avg.z) << 16) | (pack_chan(avg.w) << 24));
    }
}

// The same filter over R11G11B10f texels (Texture.h:135-185): bit-level
// unpack to float, average, clamp and cut the float's bits.
static inline vf unpack_f11(vu x) { x = x << 17; return (vf)((x & 0x0FFE0000u) + 0x38000000u); }
static inline vf unpack_f10(vu x) { x = x << 18; return (vf)((x & 0x0FFC0000u) + 0x38000000u); }
static inline vu pack_f(vf x, float hi, int shift, uint32_t mask, uint32_t bias) {
    x = x > (1.0f / (1 << 15)) ? x : (1.0f / (1 << 15));
    x = x < hi ? x : hi;
    return (((vu)x >> shift) & mask) - bias;
}

extern "C" void probe_mip_r11g11b10(const uint32_t* c00, const uint32_t* c10, const uint32_t* c01,
                                    const uint32_t* c11, uint32_t* out, int n) {
    for (int i = 0; i < n; i += 16) {
        vu p[4] = { ldu(c00 + i), ldu(c10 + i), ldu(c01 + i), ldu(c11 + i) };
        vf r = unpack_f11(p[0] >> 21), g = unpack_f11(p[0] >> 10), b = unpack_f10(p[0]);
        for (int k = 1; k < 4; k++) {
            r = r + unpack_f11(p[k] >> 21);
            g = g + unpack_f11(p[k] >> 10);
            b = b + unpack_f10(p[k]);
        }
        r = r * 0.25f; g = g * 0.25f; b = b * 0.25f;
        stu(out + i, (pack_f(r, 130048.0f, 17, 0x3FFF, 0x1C00) << 21) | (pack_f(g, 130048.0f, 17, 0x3FFF, 0x1C00) << 10) |
                         pack_f(b, 129024.0f, 18, 0x1FFF, 0x0E00));
    }
}
