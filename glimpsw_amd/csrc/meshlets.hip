// Host-side import helpers: meshlet building for indexed triangle meshes and
// the vertex attribute packing of Scene::ImportGltf (Scene.cpp:199-279).
//
// The reference builds meshlets with meshoptimizer v1.1 (meshopt_buildMeshlets
// with 64 vertices, 128 triangles and cone weight 0.25, then
// meshopt_optimizeMeshlet and meshopt_computeMeshletBounds). That library is
// not in this image; this file has its own greedy builder with the same
// limits and output layout (meshlet ranges into one vertex-index array and
// one array of u8 triangle corners), and its own bounding spheres. Meshlet
// partitions therefore differ from the reference's; every triangle of the
// input appears in exactly one meshlet.
//
// Greedy growth: a meshlet starts from a seed triangle and repeatedly takes,
// among the unused triangles that share a vertex with it, the one adding
// the fewest new vertices (ties: nearest centroid to the meshlet's, then the
// lowest index); when no neighbour fits, the lowest unused triangle if its
// three vertices fit. A meshlet closes only at the triangle limit or with
// more than max_vertices - 3 vertices, so each closed meshlet holds at least
// min(max_triangles, (max_vertices - 2) / 3) triangles. The next seed is a
// neighbour of the meshlet just closed (keeping consecutive meshlets close),
// else the lowest unused triangle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "glimp.h"

namespace {

// 128 (A, -B) pairs of the piecewise-linear VRCP14 mantissa table, as the
// device uses them (device_math.h): entry j of a run is (A - B * j) >> 12.
const uint32_t kPairs[] = {
#include "approx14_pwl.inc"
};

float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
uint32_t float_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// VRCP14PS on the host (the reference's approx_rcp, SIMD.h:294): table
// mantissa from the pairs, exponent 253 - e (254 - e for powers of two),
// zeros to infinities, infinities to zeros, NaN quietened, denormal inputs
// normalised, results below the normal range flushed to denormals or zero.
float rcp14(float xf) {
    const uint32_t x = float_to_bits(xf);
    const uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) return bits_to_float(m ? (x | 0x400000u) : s);
    if (e == 0) {
        if (m == 0) return bits_to_float(s | 0x7F800000u);
        const int lz = __builtin_clz(m) - 8;
        m <<= lz;
        e = 1 - lz;
        m &= 0x7FFFFFu;
    }
    int re;
    uint32_t mant;
    if (m == 0) {
        re = 254 - e;
        mant = 0;
    } else {
        re = 253 - e;
        const uint32_t idx = m >> 7;
        const uint32_t a = kPairs[2 * (idx >> 10)], nb = kPairs[2 * (idx >> 10) + 1];
        mant = ((a + nb * (idx & 1023)) >> 12) << 7;    // u32 arithmetic wraps as the device's
    }
    if (re >= 255) return bits_to_float(s | 0x7F800000u);
    if (re <= 0) {
        const int sh = 1 - re;
        if (sh > 24) return bits_to_float(s);
        return bits_to_float(s | ((mant | 0x800000u) >> sh));
    }
    return bits_to_float(s | ((uint32_t)re << 23) | mant);
}

// texutil::MapOctahedron (Texture.h:282-288) as the reference's optimised
// build evaluates it: 1/(|x|+|y|+|z|) from VRCP14PS alone, max(-z * w, 0)
// zero unless positive, (c * 0.5 + 0.5) one FMA.
void map_octahedron(const float n[3], float& u, float& v) {
    const float w = rcp14((std::fabs(n[0]) + std::fabs(n[1])) + std::fabs(n[2]));
    const float mz = -n[2] * w;
    const float t = mz > 0.0f ? mz : 0.0f;
    const float sx = bits_to_float(float_to_bits(t) ^ (float_to_bits(n[0]) & 0x80000000u));
    const float sy = bits_to_float(float_to_bits(t) ^ (float_to_bits(n[1]) & 0x80000000u));
    u = std::fma(std::fma(n[0], w, sx), 0.5f, 0.5f);
    v = std::fma(std::fma(n[1], w, sy), 0.5f, 0.5f);
}

// glm::packUnorm4x8 (glm 1.0.1): round(clamp(c, 0, 1) * 255) per byte, with
// round() halfway cases away from zero.
uint32_t pack_unorm4x8(const float c[4]) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) {
        const float x = std::min(std::max(c[k], 0.0f), 1.0f);
        r |= (uint32_t)std::round(x * 255.0f) << (8 * k);
    }
    return r;
}

struct Vec3 { double x, y, z; };

Vec3 vertex(const float* positions, size_t stride, uint32_t i) {
    const float* p = (const float*)((const char*)positions + (size_t)i * stride);
    return { p[0], p[1], p[2] };
}

}  // namespace

extern "C" {

size_t glimp_build_meshlets_bound(size_t index_count, size_t max_vertices, size_t max_triangles) {
    if (max_vertices < 3 || max_triangles < 1) return 0;
    const size_t tris = index_count / 3;
    // every meshlet but the last holds at least max_vertices / 3 triangles
    // or max_triangles of them
    const size_t per = std::max<size_t>(1, std::min(max_triangles, (max_vertices - 2) / 3));
    return (tris + per - 1) / per;
}

int glimp_build_meshlets(const uint32_t* indices, size_t index_count, const float* positions, size_t vertex_count,
                         size_t vertex_stride, uint32_t max_vertices, uint32_t max_triangles,
                         glimp_meshlet_range* meshlets, uint32_t* meshlet_vertices, uint8_t* meshlet_triangles,
                         size_t* meshlet_count) {
    if (!indices || !positions || !meshlets || !meshlet_vertices || !meshlet_triangles || !meshlet_count)
        return GLIMP_E_INVALID;
    if (index_count % 3 || max_vertices < 3 || max_vertices > 256 || max_triangles < 1 || max_triangles > 512 ||
        vertex_stride < 12)
        return GLIMP_E_INVALID;
    const size_t T = index_count / 3;
    for (size_t i = 0; i < index_count; i++)
        if (indices[i] >= vertex_count) return GLIMP_E_INVALID;
    // vertex -> triangles (CSR)
    std::vector<uint32_t> start(vertex_count + 1, 0), adj(index_count);
    for (size_t i = 0; i < index_count; i++) start[indices[i] + 1]++;
    for (size_t v = 0; v < vertex_count; v++) start[v + 1] += start[v];
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < index_count; i++) adj[fill[indices[i]]++] = (uint32_t)(i / 3);
    }
    std::vector<Vec3> centroid(T);
    for (size_t t = 0; t < T; t++) {
        Vec3 a = vertex(positions, vertex_stride, indices[3 * t]), b = vertex(positions, vertex_stride, indices[3 * t + 1]),
             c = vertex(positions, vertex_stride, indices[3 * t + 2]);
        centroid[t] = { (a.x + b.x + c.x) / 3, (a.y + b.y + c.y) / 3, (a.z + b.z + c.z) / 3 };
    }
    std::vector<uint8_t> used(T, 0);
    std::vector<int32_t> local(vertex_count, -1);   // vertex -> slot in the open meshlet
    std::vector<uint32_t> verts, tris, cand;
    size_t nm = 0, vout = 0, tout = 0, lowest = 0, remaining = T;
    int64_t next_seed = -1;
    while (remaining) {
        int64_t seed = next_seed;
        if (seed < 0 || used[seed]) {
            while (used[lowest]) lowest++;
            seed = (int64_t)lowest;
        }
        verts.clear();
        tris.clear();
        cand.clear();
        Vec3 sum = { 0, 0, 0 };
        auto add = [&](uint32_t t) {
            used[t] = 1;
            remaining--;
            tris.push_back(t);
            sum.x += centroid[t].x; sum.y += centroid[t].y; sum.z += centroid[t].z;
            for (int k = 0; k < 3; k++) {
                const uint32_t v = indices[3 * t + k];
                if (local[v] >= 0) continue;
                local[v] = (int32_t)verts.size();
                verts.push_back(v);
                for (uint32_t a = start[v]; a < start[v + 1]; a++)
                    if (!used[adj[a]]) cand.push_back(adj[a]);
            }
        };
        add((uint32_t)seed);
        while (tris.size() < max_triangles) {
            int64_t best = -1;
            int bestNew = 4;
            double bestDist = 0;
            size_t keep = 0;
            const double n = (double)tris.size();
            const Vec3 c = { sum.x / n, sum.y / n, sum.z / n };
            for (size_t i = 0; i < cand.size(); i++) {
                const uint32_t t = cand[i];
                if (used[t]) continue;
                cand[keep++] = t;     // drop used entries as we go
                int fresh = 0;
                for (int k = 0; k < 3; k++) fresh += local[indices[3 * t + k]] < 0 ? 1 : 0;
                // a corner listed twice (degenerate triangle) is counted twice; harmless
                if (verts.size() + fresh > max_vertices) continue;
                const double dx = centroid[t].x - c.x, dy = centroid[t].y - c.y, dz = centroid[t].z - c.z;
                const double d = dx * dx + dy * dy + dz * dz;
                if (fresh < bestNew || (fresh == bestNew && (d < bestDist || (d == bestDist && t < best)))) {
                    best = t;
                    bestNew = fresh;
                    bestDist = d;
                }
            }
            cand.resize(keep);
            if (best < 0) {
                if (!remaining || verts.size() + 3 > max_vertices) break;
                while (used[lowest]) lowest++;
                best = (int64_t)lowest;
            }
            add((uint32_t)best);
        }
        // emit
        glimp_meshlet_range& m = meshlets[nm++];
        m.vertex_offset = (uint32_t)vout;
        m.triangle_offset = (uint32_t)tout;
        m.vertex_count = (uint32_t)verts.size();
        m.triangle_count = (uint32_t)tris.size();
        for (uint32_t v : verts) meshlet_vertices[vout++] = v;
        for (uint32_t t : tris)
            for (int k = 0; k < 3; k++) meshlet_triangles[tout++] = (uint8_t)local[indices[3 * t + k]];
        next_seed = -1;
        for (uint32_t t : cand)
            if (!used[t]) { next_seed = t; break; }
        for (uint32_t v : verts) local[v] = -1;
    }
    *meshlet_count = nm;
    return GLIMP_OK;
}

int glimp_meshlet_bounds(const uint32_t* meshlet_vertices, uint32_t vertex_count, const float* positions,
                         size_t vertex_stride, float sphere[4]) {
    if (!meshlet_vertices || !positions || !sphere || vertex_count == 0) return GLIMP_E_INVALID;
    // Ritter's sphere: start from the farthest pair found from the first
    // vertex, grow over the rest, then widen to cover every vertex in float
    auto P = [&](uint32_t i) { return vertex(positions, vertex_stride, meshlet_vertices[i]); };
    auto d2 = [](Vec3 a, Vec3 b) { return (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z); };
    Vec3 a = P(0), b = a;
    for (uint32_t i = 0; i < vertex_count; i++) if (d2(P(i), a) > d2(b, a)) b = P(i);
    Vec3 c2 = b;
    for (uint32_t i = 0; i < vertex_count; i++) if (d2(P(i), b) > d2(c2, b)) c2 = P(i);
    Vec3 ctr = { (b.x + c2.x) / 2, (b.y + c2.y) / 2, (b.z + c2.z) / 2 };
    double r = std::sqrt(d2(b, c2)) / 2;
    for (uint32_t i = 0; i < vertex_count; i++) {
        const Vec3 p = P(i);
        const double d = std::sqrt(d2(p, ctr));
        if (d > r) {
            const double nr = (r + d) / 2, k = (nr - r) / d;
            ctr = { ctr.x + (p.x - ctr.x) * k, ctr.y + (p.y - ctr.y) * k, ctr.z + (p.z - ctr.z) * k };
            r = nr;
        }
    }
    const float cx = (float)ctr.x, cy = (float)ctr.y, cz = (float)ctr.z;
    double rr = r;
    for (uint32_t i = 0; i < vertex_count; i++) {
        const Vec3 p = P(i);
        rr = std::max(rr, std::sqrt(d2(p, { cx, cy, cz })));
    }
    sphere[0] = cx;
    sphere[1] = cy;
    sphere[2] = cz;
    sphere[3] = std::nextafter((float)rr, INFINITY);
    return GLIMP_OK;
}

int glimp_pack_normal_tangents(const float* normals, const float* tangents, size_t count, uint32_t* packed) {
    if (!normals || !packed) return GLIMP_E_INVALID;
    for (size_t i = 0; i < count; i++) {
        float c[4];
        map_octahedron(normals + 3 * i, c[0], c[1]);
        if (tangents) {
            map_octahedron(tangents + 4 * i, c[2], c[3]);
        } else {
            c[2] = c[3] = 0.0f;
        }
        packed[i] = pack_unorm4x8(c);
    }
    return GLIMP_OK;
}

}  // extern "C"
