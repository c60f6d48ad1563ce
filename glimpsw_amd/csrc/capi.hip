// C ABI of the glimp library (include/glimp.h): resource management and
// draw/read entry points over the kernels in glimp_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "glimp.h"
#include "internal.h"
#include "resolve.h"

// 131072 u16: VRCP14 then VRSQRT14 mantissa tables, generated at build time
// from glimpsw_amd/data/approx14.bin.
static const uint16_t kApproxTables[] = {
#include "approx14_tables.inc"
};
static_assert(sizeof(kApproxTables) == 131072 * 2, "approximation tables");

using namespace glimp;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(GLIMP_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e));  \
    } while (0)

template <typename T>
hipError_t dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

template <typename T>
hipError_t upload(T** dst, const T* src, size_t count) {
    hipError_t e = dev_alloc(dst, count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

struct glimp_context {
    int device = 0;
    uint16_t* tables = nullptr;
    Counters* counters = nullptr;
    DrawScratch* scratch = nullptr;
    uint32_t* visible = nullptr;
    TriRecord* records = nullptr;
    size_t records_cap = 0;
    ClipEntry* clip = nullptr;
    size_t clip_cap = 0;
    uint16_t* bitmap = nullptr;
    size_t bitmap_cap = 0;
    uint32_t* bin_counts = nullptr;
    size_t bin_counts_cap = 0;
    uint32_t* bin_lists = nullptr;
    size_t bin_lists_cap = 0;
    uint32_t* spill = nullptr;
    size_t spill_cap = 0;
    uint32_t* meshlet_counts = nullptr;
    size_t meshlet_counts_cap = 0;
    Profiler prof;

    ApproxTables approx() const { return ApproxTables{ tables, tables + 65536 }; }
};

struct glimp_texture {
    glimp_context* ctx = nullptr;
    DevTexture hdr = {};
    uint32_t* data = nullptr;
};

struct glimp_scene {
    glimp_context* ctx = nullptr;
    DevMeshlets m = {};
    std::vector<void*> buffers;
    DevMaterial* materials = nullptr;
    uint32_t num_materials = 0;
    DevTexture* textures = nullptr;
    uint32_t num_textures = 0;
};

struct glimp_framebuffer {
    glimp_context* ctx = nullptr;
    uint32_t width = 0, height = 0;
    unsigned long long* keys = nullptr;
    uint32_t clear_color = 0;
    float clear_depth = 0.0f;
    bool pending_clear = false;        // cleared, keys not yet written
    uint32_t next_seq = 0;
    std::vector<DrawEntry> draws;
    DrawEntry* d_draws = nullptr;
    size_t d_draws_cap = 0;
    uint32_t* color = nullptr;         // resolved RGBA8, row-major
};

namespace {

hipError_t flush_clear(glimp_framebuffer* f, hipStream_t s) {
    if (!f->pending_clear) return hipSuccess;
    uint32_t db;
    memcpy(&db, &f->clear_depth, 4);
    hipError_t e = launch_clear(f->keys, f->width * f->height, (unsigned long long)db << 32, s);
    if (e == hipSuccess) f->pending_clear = false;
    return e;
}

template <typename T>
int grow_one(T*& buf, size_t& cap, size_t need) {
    if (need <= cap) return GLIMP_OK;
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    HIP_TRY(dev_alloc(&buf, need));
    cap = need;
    return GLIMP_OK;
}

int grow(glimp_context* c, size_t records, size_t clip, size_t bitmap) {
    if (records > c->records_cap) {
        if (c->records) (void)hipFree(c->records);
        c->records = nullptr;
        c->records_cap = 0;
        HIP_TRY(dev_alloc(&c->records, records));
        c->records_cap = records;
    }
    if (clip > c->clip_cap) {
        if (c->clip) (void)hipFree(c->clip);
        c->clip = nullptr;
        c->clip_cap = 0;
        HIP_TRY(dev_alloc(&c->clip, clip));
        c->clip_cap = clip;
    }
    if (bitmap > c->bitmap_cap) {
        if (c->bitmap) (void)hipFree(c->bitmap);
        c->bitmap = nullptr;
        c->bitmap_cap = 0;
        HIP_TRY(dev_alloc(&c->bitmap, bitmap));
        c->bitmap_cap = bitmap;
    }
    return GLIMP_OK;
}

}  // namespace

extern "C" {

const char* glimp_last_error(void) { return g_error.c_str(); }

int glimp_create(glimp_context** out, int device) {
    if (!out) return fail(GLIMP_E_INVALID, "glimp_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(GLIMP_E_INVALID, "glimp_create: no such device");
    HIP_TRY(hipSetDevice(device));
    glimp_context* c = new (std::nothrow) glimp_context();
    if (!c) return fail(GLIMP_E_NOMEM, "glimp_create: out of host memory");
    c->device = device;
    int rc = GLIMP_OK;
    hipError_t e = upload(&c->tables, kApproxTables, 131072);
    if (e == hipSuccess) e = dev_alloc(&c->counters, 1);
    if (e == hipSuccess) e = hipMemset(c->counters, 0, sizeof(Counters));
    if (e == hipSuccess) e = dev_alloc(&c->scratch, 1);
    if (e == hipSuccess) e = hipMemset(c->scratch, 0, sizeof(DrawScratch));
    if (e == hipSuccess) e = dev_alloc(&c->visible, 1);
    if (e != hipSuccess) rc = fail(GLIMP_E_HIP, std::string("glimp_create: ") + hipGetErrorString(e));
    if (rc != GLIMP_OK) {
        glimp_destroy(c);
        return rc;
    }
    *out = c;
    return GLIMP_OK;
}

void glimp_destroy(glimp_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    void* ptrs[] = { c->tables, c->counters, c->scratch, c->visible, c->records, c->clip, c->bitmap,
                     c->bin_counts, c->bin_lists, c->spill, c->meshlet_counts };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete c;
}

int glimp_scene_create(glimp_context* c, const glimp_meshlets_desc* d, glimp_scene** out) {
    if (!c || !d || !out) return fail(GLIMP_E_INVALID, "glimp_scene_create: NULL argument");
    *out = nullptr;
    if (!d->bounds || !d->num_vertices || !d->num_triangles || !d->material_id || !d->positions ||
        !d->texcoords || !d->normal_tangents || !d->indices)
        return fail(GLIMP_E_INVALID, "glimp_scene_create: missing array");
    for (uint32_t i = 0; i < d->count; i++) {
        if (d->num_vertices[i] > GLIMP_MAX_VERTICES || d->num_triangles[i] > GLIMP_MAX_PRIMS)
            return fail(GLIMP_E_LIMIT, "glimp_scene_create: meshlet exceeds 64 vertices / 128 triangles");
    }
    HIP_TRY(hipSetDevice(c->device));
    glimp_scene* s = new (std::nothrow) glimp_scene();
    if (!s) return fail(GLIMP_E_NOMEM, "glimp_scene_create: out of host memory");
    s->ctx = c;
    size_t n = d->count;
    s->m.count = d->count;
    hipError_t e = hipSuccess;
    auto up = [&](auto** dst, const auto* src, size_t cnt) {
        if (e != hipSuccess) return;
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        T* p = nullptr;
        e = upload(&p, src, cnt);
        if (e == hipSuccess) {
            *dst = p;
            s->buffers.push_back(p);
        }
    };
    std::vector<uint64_t> zeros;
    const uint64_t* hand = d->tangent_handedness;
    if (!hand) {
        zeros.assign(n, 0);
        hand = zeros.data();
    }
    const float4* bounds = reinterpret_cast<const float4*>(d->bounds);
    up((float4**)&s->m.bounds, bounds, n);
    up((uint8_t**)&s->m.num_vertices, d->num_vertices, n);
    up((uint8_t**)&s->m.num_triangles, d->num_triangles, n);
    up((uint32_t**)&s->m.material_id, d->material_id, n);
    up((uint64_t**)&s->m.tangent_handedness, hand, n);
    up((float**)&s->m.positions, d->positions, n * 192);
    up((uint32_t**)&s->m.texcoords, d->texcoords, n * 64);
    up((uint32_t**)&s->m.normal_tangents, d->normal_tangents, n * 64);
    up((uint8_t**)&s->m.indices, d->indices, n * 384);
    if (e != hipSuccess) {
        glimp_scene_destroy(s);
        return fail(GLIMP_E_HIP, std::string("glimp_scene_create: ") + hipGetErrorString(e));
    }
    *out = s;
    return GLIMP_OK;
}

void glimp_scene_destroy(glimp_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    for (void* p : s->buffers) (void)hipFree(p);
    if (s->materials) (void)hipFree(s->materials);
    if (s->textures) (void)hipFree(s->textures);
    delete s;
}

int glimp_texture_create(glimp_context* c, const glimp_texture_desc* d, glimp_texture** out) {
    if (!c || !d || !out || !d->texels) return fail(GLIMP_E_INVALID, "glimp_texture_create: NULL argument");
    *out = nullptr;
    auto pow2 = [](uint32_t v) { return v && !(v & (v - 1)); };
    if (!pow2(d->width) || !pow2(d->height) || d->mip_levels < 1 || d->mip_levels > 16 || d->num_layers < 1)
        return fail(GLIMP_E_INVALID, "glimp_texture_create: sizes must be powers of two, 1-16 mips");
    HIP_TRY(hipSetDevice(c->device));
    glimp_texture* t = new (std::nothrow) glimp_texture();
    if (!t) return fail(GLIMP_E_NOMEM, "glimp_texture_create: out of host memory");
    t->ctx = c;
    size_t n = (size_t)d->layer_stride * d->num_layers;
    // 64 texels of slack: bilinear reads the right neighbour of the last
    // column at offset +8 even when its weight is zero (Texture.h:527-531)
    hipError_t e = dev_alloc(&t->data, n + 64);
    if (e == hipSuccess) e = hipMemset(t->data, 0, (n + 64) * 4);
    if (e == hipSuccess) e = hipMemcpy(t->data, d->texels, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (t->data) (void)hipFree(t->data);
        delete t;
        return fail(GLIMP_E_HIP, std::string("glimp_texture_create: ") + hipGetErrorString(e));
    }
    DevTexture& h = t->hdr;
    h.width = d->width;
    h.height = d->height;
    h.mip_levels = d->mip_levels;
    h.num_layers = d->num_layers;
    h.row_shift = d->row_shift;
    h.layer_stride = d->layer_stride;
    h.scale_u = (float)d->width;
    h.scale_v = (float)d->height;
    h.mask_lerp_u = (int32_t)(d->width << 8) - 1;
    h.mask_lerp_v = (int32_t)(d->height << 8) - 1;
    h.scale_lerp_u = (float)(h.mask_lerp_u + 1);
    h.scale_lerp_v = (float)(h.mask_lerp_v + 1);
    memcpy(h.mip_offsets, d->mip_offsets, sizeof(h.mip_offsets));
    h.data = t->data;
    *out = t;
    return GLIMP_OK;
}

void glimp_texture_destroy(glimp_texture* t) {
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    if (t->data) (void)hipFree(t->data);
    delete t;
}

int glimp_scene_set_materials(glimp_scene* s, const glimp_material* mats, uint32_t count,
                              glimp_texture* const* textures, uint32_t num_textures) {
    if (!s || (count && !mats) || (num_textures && !textures))
        return fail(GLIMP_E_INVALID, "glimp_scene_set_materials: NULL argument");
    HIP_TRY(hipSetDevice(s->ctx->device));
    std::vector<DevMaterial> dm(count);
    for (uint32_t i = 0; i < count; i++) {
        if (mats[i].texture >= (int32_t)num_textures)
            return fail(GLIMP_E_INVALID, "glimp_scene_set_materials: texture index out of range");
        dm[i].texture = mats[i].texture;
        dm[i].cull_mode = mats[i].double_sided ? 0 : 1;
        dm[i].fs_id = mats[i].alpha_cutoff < 255 ? 1 : 0;
        dm[i].alpha_cutoff = mats[i].alpha_cutoff;
        dm[i].pad = 0;
    }
    std::vector<DevTexture> dt(num_textures);
    for (uint32_t i = 0; i < num_textures; i++) {
        if (!textures[i]) return fail(GLIMP_E_INVALID, "glimp_scene_set_materials: NULL texture");
        dt[i] = textures[i]->hdr;
    }
    if (s->materials) (void)hipFree(s->materials);
    if (s->textures) (void)hipFree(s->textures);
    s->materials = nullptr;
    s->textures = nullptr;
    HIP_TRY(upload(&s->materials, dm.data(), dm.size()));
    HIP_TRY(upload(&s->textures, dt.data(), dt.size()));
    s->num_materials = count;
    s->num_textures = num_textures;
    return GLIMP_OK;
}

int glimp_framebuffer_create(glimp_context* c, uint32_t width, uint32_t height, glimp_framebuffer** out) {
    if (!c || !out) return fail(GLIMP_E_INVALID, "glimp_framebuffer_create: NULL argument");
    *out = nullptr;
    if (width == 0 || height == 0 || width % 4 || height % 4)
        return fail(GLIMP_E_INVALID, "glimp_framebuffer_create: width and height must be non-zero multiples of 4");
    if (width > GLIMP_MAX_RENDER_SIZE || height > GLIMP_MAX_RENDER_SIZE)
        return fail(GLIMP_E_LIMIT, "glimp_framebuffer_create: larger than GLIMP_MAX_RENDER_SIZE");
    HIP_TRY(hipSetDevice(c->device));
    glimp_framebuffer* f = new (std::nothrow) glimp_framebuffer();
    if (!f) return fail(GLIMP_E_NOMEM, "glimp_framebuffer_create: out of host memory");
    f->ctx = c;
    f->width = width;
    f->height = height;
    hipError_t e = dev_alloc(&f->keys, (size_t)width * height);
    if (e == hipSuccess) e = hipMemset(f->keys, 0, (size_t)width * height * 8);
    if (e != hipSuccess) {
        delete f;
        return fail(GLIMP_E_HIP, std::string("glimp_framebuffer_create: ") + hipGetErrorString(e));
    }
    *out = f;
    return GLIMP_OK;
}

void glimp_framebuffer_destroy(glimp_framebuffer* f) {
    if (!f) return;
    (void)hipSetDevice(f->ctx->device);
    if (f->keys) (void)hipFree(f->keys);
    if (f->d_draws) (void)hipFree(f->d_draws);
    if (f->color) (void)hipFree(f->color);
    delete f;
}

int glimp_framebuffer_clear(glimp_framebuffer* f, uint32_t color, float depth, void* stream) {
    if (!f) return fail(GLIMP_E_INVALID, "glimp_framebuffer_clear: NULL framebuffer");
    if (!(depth >= 0.0f)) return fail(GLIMP_E_INVALID, "glimp_framebuffer_clear: depth must be >= 0");
    if (depth == 0.0f) depth = 0.0f;  // -0 -> +0
    HIP_TRY(hipSetDevice(f->ctx->device));
    unsigned long long key;
    uint32_t db;
    memcpy(&db, &depth, 4);
    key = (unsigned long long)db << 32;
    (void)key;
    (void)stream;
    // The keys are written by the first draw (each bin starts from the clear
    // key) or, if no binned draw follows, by flush_clear before use.
    f->clear_color = color;
    f->clear_depth = depth;
    f->next_seq = 0;
    f->draws.clear();
    f->pending_clear = true;
    return GLIMP_OK;
}

int glimp_framebuffer_keys(glimp_framebuffer* f, void** keys, size_t* bytes) {
    if (!f || !keys) return fail(GLIMP_E_INVALID, "glimp_framebuffer_keys: NULL argument");
    HIP_TRY(hipSetDevice(f->ctx->device));
    HIP_TRY(flush_clear(f, nullptr));
    *keys = f->keys;
    if (bytes) *bytes = (size_t)f->width * f->height * 8;
    return GLIMP_OK;
}

namespace {

// Uploads the framebuffer's draw table (seq ranges -> meshlet offsets).
int upload_draws(glimp_framebuffer* f, hipStream_t s, uint32_t* count) {
    std::vector<DrawEntry> draws = f->draws;
    if (draws.empty()) draws.push_back(DrawEntry{ 0, 0, 0, 0 });
    if (draws.size() > f->d_draws_cap) {
        if (f->d_draws) (void)hipFree(f->d_draws);
        f->d_draws = nullptr;
        f->d_draws_cap = 0;
        HIP_TRY(dev_alloc(&f->d_draws, draws.size()));
        f->d_draws_cap = draws.size();
    }
    HIP_TRY(hipMemcpyAsync(f->d_draws, draws.data(), draws.size() * sizeof(DrawEntry), hipMemcpyHostToDevice, s));
    // the host vector must outlive the async copy
    HIP_TRY(hipStreamSynchronize(s));
    *count = (uint32_t)draws.size();
    return GLIMP_OK;
}

}  // namespace

int glimp_framebuffer_read(glimp_framebuffer* f, uint32_t* ids, float* depth, int linear, void* stream) {
    if (!f) return fail(GLIMP_E_INVALID, "glimp_framebuffer_read: NULL framebuffer");
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    size_t n = (size_t)f->width * f->height;
    HIP_TRY(flush_clear(f, s));
    uint32_t ndraws = 0;
    int rc = upload_draws(f, s, &ndraws);
    if (rc != GLIMP_OK) return rc;
    uint32_t* d_ids = nullptr;
    float* d_depth = nullptr;
    HIP_TRY(dev_alloc(&d_ids, n));
    hipError_t e = dev_alloc(&d_depth, n);
    uint32_t db;
    memcpy(&db, &f->clear_depth, 4);
    if (e == hipSuccess)
        e = launch_decode(f->keys, (int)f->width, (int)f->height, (unsigned long long)db << 32, f->clear_color,
                          f->d_draws, ndraws, d_ids, d_depth, linear, s);
    if (e == hipSuccess && ids) e = hipMemcpyAsync(ids, d_ids, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, d_depth, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_ids);
    if (d_depth) (void)hipFree(d_depth);
    if (e != hipSuccess) return fail(GLIMP_E_HIP, std::string("glimp_framebuffer_read: ") + hipGetErrorString(e));
    return GLIMP_OK;
}

int glimp_draw_meshlets(glimp_context* c, glimp_framebuffer* f, glimp_scene* s, const glimp_draw_desc* d,
                        void* stream) {
    if (!c || !f || !s || !d) return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: NULL argument");
    if (d->pipeline != GLIMP_PIPELINE_VISBUFFER)
        return fail(GLIMP_E_UNSUPPORTED, "glimp_draw_meshlets: unknown pipeline");
    if ((uint64_t)d->meshlet_offset + d->meshlet_count > s->m.count)
        return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: meshlet range outside the scene");
    if ((uint64_t)f->next_seq + d->meshlet_count > GLIMP_MAX_DRAWN_MESHLETS)
        return fail(GLIMP_E_LIMIT, "glimp_draw_meshlets: more than GLIMP_MAX_DRAWN_MESHLETS meshlets since the last clear");
    if (d->meshlet_count == 0) return GLIMP_OK;
    HIP_TRY(hipSetDevice(c->device));
    size_t cnt = d->meshlet_count;
    // worst case: every triangle clipped, each into 7 fan triangles
    int rc = grow(c, cnt * 128 * 8, cnt * 128, (cnt + 15) / 16);
    if (rc == GLIMP_OK) rc = grow_one(c->meshlet_counts, c->meshlet_counts_cap, cnt);
    if (rc != GLIMP_OK) return rc;

    DrawParams p;
    memset(&p, 0, sizeof(p));
    p.m = s->m;
    p.materials = s->materials;
    p.num_materials = s->num_materials;
    p.meshlet_offset = d->meshlet_offset;
    p.meshlet_count = d->meshlet_count;
    p.seq_base = f->next_seq;
    memcpy(p.clip, d->object_to_clip, sizeof(p.clip));
    p.use_cull = d->frustum_planes != nullptr;
    if (p.use_cull) memcpy(p.planes, d->frustum_planes, sizeof(p.planes));
    p.width = (int)f->width;
    p.height = (int)f->height;
    bool guard = !(d->flags & GLIMP_DRAW_NO_GUARDBAND);
    p.guard_x = guard ? (float)GLIMP_MAX_RENDER_SIZE / (float)f->width : 1.0f;
    p.guard_y = guard ? (float)GLIMP_MAX_RENDER_SIZE / (float)f->height : 1.0f;
    p.clipping = !(d->flags & GLIMP_DRAW_NO_CLIPPING);
    p.flags = d->flags;
    p.tables = c->approx();
    p.cull_bitmap = c->bitmap;
    p.records = c->records;
    p.meshlet_counts = c->meshlet_counts;
    p.clip_record_base = d->meshlet_count * 128;
    p.clip_record_capacity = (uint32_t)std::min<size_t>(c->records_cap - (size_t)d->meshlet_count * 128, 0x7FFFFFFFu);
    p.clip_queue = c->clip;
    p.clip_capacity = (uint32_t)std::min<size_t>(c->clip_cap, 0xFFFFFFFFu);
    p.scratch = c->scratch;
    p.counters = c->counters;
    p.keys = f->keys;
    p.clear_depth = f->clear_depth;
    hipStream_t st = (hipStream_t)stream;
    if (d->flags & GLIMP_DRAW_DIRECT) {
        HIP_TRY(flush_clear(f, st));
        HIP_TRY(launch_draw_direct(p, st, &c->prof));
    } else {
        uint32_t bs = d->bin_size ? d->bin_size : 16;
        if (bs != 8 && bs != 16 && bs != 32 && bs != 64)
            return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: bin_size must be 8, 16, 32 or 64");
        uint32_t cap = d->bin_capacity ? d->bin_capacity : 1024;
        BinParams& B = p.bins;
        B.bin_size = bs;
        B.bin_shift = (uint32_t)__builtin_ctz(bs);
        B.bins_x = (f->width + bs - 1) / bs;
        B.bins_y = (f->height + bs - 1) / bs;
        B.bin_capacity = cap;
        size_t nb = (size_t)B.bins_x * B.bins_y;
        rc = grow_one(c->bin_counts, c->bin_counts_cap, nb);
        if (rc == GLIMP_OK) rc = grow_one(c->bin_lists, c->bin_lists_cap, nb * cap);
        if (rc == GLIMP_OK) rc = grow_one(c->spill, c->spill_cap, c->records_cap);
        if (rc != GLIMP_OK) return rc;
        B.bin_counts = c->bin_counts;
        B.bin_lists = c->bin_lists;
        B.spill = c->spill;
        B.fresh = f->pending_clear ? 1 : 0;
        HIP_TRY(launch_draw_binned(p, st, &c->prof));
        f->pending_clear = false;
    }
    f->draws.push_back(DrawEntry{ f->next_seq, d->meshlet_count, d->meshlet_offset, 0 });
    f->next_seq += d->meshlet_count;
    return GLIMP_OK;
}

int glimp_cull_meshlets(glimp_context* c, glimp_scene* s, uint32_t first, uint32_t count, const float planes[20],
                        uint16_t* bitmap, uint32_t* visible, void* stream) {
    if (!c || !s || !planes) return fail(GLIMP_E_INVALID, "glimp_cull_meshlets: NULL argument");
    if ((uint64_t)first + count > s->m.count) return fail(GLIMP_E_INVALID, "glimp_cull_meshlets: range outside the scene");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    size_t words = (count + 15) / 16;
    int rc = grow(c, 0, 0, words);
    if (rc != GLIMP_OK) return rc;
    HIP_TRY(hipMemsetAsync(c->visible, 0, 4, st));
    if (count) HIP_TRY(launch_cull(s->m, first, count, planes, c->bitmap, c->visible, st));
    if (bitmap && words) HIP_TRY(hipMemcpyAsync(bitmap, c->bitmap, words * 2, hipMemcpyDeviceToHost, st));
    uint32_t v = 0;
    HIP_TRY(hipMemcpyAsync(&v, c->visible, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (visible) *visible = v;
    return GLIMP_OK;
}

int glimp_resolve(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_resolve_desc* d, void* stream) {
    if (!c || !f || !sc || !d) return fail(GLIMP_E_INVALID, "glimp_resolve: NULL argument");
    if (d->mode > GLIMP_RESOLVE_LIT) return fail(GLIMP_E_INVALID, "glimp_resolve: unknown mode");
    if (d->num_lights > GLIMP_MAX_LIGHTS || (d->num_lights && !d->lights))
        return fail(GLIMP_E_LIMIT, "glimp_resolve: at most GLIMP_MAX_LIGHTS lights");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(flush_clear(f, s));
    if (!f->color) HIP_TRY(dev_alloc(&f->color, (size_t)f->width * f->height));
    ResolveParams p;
    memset(&p, 0, sizeof(p));
    p.m = sc->m;
    p.materials = sc->materials;
    p.num_materials = sc->num_materials;
    p.textures = sc->textures;
    p.tables = c->approx();
    p.keys = f->keys;
    p.width = (int)f->width;
    p.height = (int)f->height;
    uint32_t db;
    memcpy(&db, &f->clear_depth, 4);
    p.clear_key = (unsigned long long)db << 32;
    int rc = upload_draws(f, s, &p.num_draws);
    if (rc != GLIMP_OK) return rc;
    p.draws = f->d_draws;
    memcpy(p.object_to_clip, d->object_to_clip, sizeof(p.object_to_clip));
    memcpy(p.object_to_world, d->object_to_world, sizeof(p.object_to_world));
    memcpy(p.inv_screen_proj, d->inv_screen_proj, sizeof(p.inv_screen_proj));
    memcpy(p.view_pos, d->view_pos, sizeof(p.view_pos));
    p.exposure = d->exposure;
    p.light_exposure = d->exposure * 0.001f;
    // screen UV of pixel centres (Rasterizer.h:226-237) and the derivative
    // steps of IntersectTriangle (Shading.cpp:454-455), IEEE scalar ops
    float sw = (float)f->width, sh = (float)f->height;
    p.scale_x = 2.0f / sw;
    p.scale_y = 2.0f / sh;
    p.center_x = std::fma(0.5f, p.scale_x, -1.0f);
    p.center_y = std::fma(0.5f, p.scale_y, -1.0f);
    p.kx = 2.0f / sw;
    p.ky = -2.0f / sh;
    p.num_lights = d->num_lights;
    for (uint32_t i = 0; i < d->num_lights; i++) memcpy(&p.lights[i], &d->lights[i], sizeof(ResolveLight));
    p.mode = (int)d->mode;
    p.out = f->color;
    p.linear = 1;
    HIP_TRY(launch_resolve(p, s, &c->prof));
    return GLIMP_OK;
}

int glimp_framebuffer_color(glimp_framebuffer* f, uint32_t** color) {
    if (!f || !color) return fail(GLIMP_E_INVALID, "glimp_framebuffer_color: NULL argument");
    *color = f->color;
    if (!f->color) return fail(GLIMP_E_INVALID, "glimp_framebuffer_color: nothing resolved yet");
    return GLIMP_OK;
}

int glimp_framebuffer_read_color(glimp_framebuffer* f, uint32_t* rgba, void* stream) {
    if (!f || !rgba) return fail(GLIMP_E_INVALID, "glimp_framebuffer_read_color: NULL argument");
    if (!f->color) return fail(GLIMP_E_INVALID, "glimp_framebuffer_read_color: nothing resolved yet");
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(rgba, f->color, (size_t)f->width * f->height * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GLIMP_OK;
}

int glimp_get_counters(glimp_context* c, glimp_counters* out) {
    if (!c || !out) return fail(GLIMP_E_INVALID, "glimp_get_counters: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    Counters h;
    HIP_TRY(hipMemcpy(&h, c->counters, sizeof(h), hipMemcpyDeviceToHost));
    out->triangles_processed = h.triangles_processed;
    out->triangles_rasterized = h.triangles_rasterized;
    out->triangles_clipped = h.triangles_clipped;
    out->bins_rasterized = h.bins_rasterized;
    out->fragments_covered = h.fragments_covered;
    out->fragments_passed = h.fragments_passed;
    out->setup_records = h.setup_records;
    out->bin_entries = h.bin_entries;
    return GLIMP_OK;
}

int glimp_profile_enable(glimp_context* c, int enable) {
    if (!c) return fail(GLIMP_E_INVALID, "glimp_profile_enable: NULL context");
    c->prof.enabled = enable != 0;
    return GLIMP_OK;
}

int glimp_profile_read(glimp_context* c, glimp_kernel_time* out, uint32_t capacity, uint32_t* count, int reset) {
    if (!c || !count) return fail(GLIMP_E_INVALID, "glimp_profile_read: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(c->prof.collect());
    uint32_t n = 0;
    for (const auto& kv : c->prof.stats()) {
        if (out && n < capacity) {
            memset(&out[n], 0, sizeof(out[n]));
            strncpy(out[n].name, kv.first.c_str(), sizeof(out[n].name) - 1);
            out[n].launches = kv.second.launches;
            out[n].total_ms = kv.second.total_ms;
        }
        n++;
    }
    *count = n;
    if (reset) c->prof.reset();
    return GLIMP_OK;
}

int glimp_reset_counters(glimp_context* c) {
    if (!c) return fail(GLIMP_E_INVALID, "glimp_reset_counters: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(c->counters, 0, sizeof(Counters)));
    return GLIMP_OK;
}

}  // extern "C"
