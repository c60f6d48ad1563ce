// C ABI of the glimp library (include/glimp.h): resource management and
// draw/read entry points over the kernels in glimp_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "glimp.h"
#include "internal.h"
#include "resolve.h"
#include "diff.h"

#include <rocprofiler-sdk-roctx/roctx.h>

// 128 (A, -B) pairs of the piecewise-linear VRCP14 / VRSQRT14 mantissa
// tables (device_math.h), generated at build time from
// glimpsw_amd/data/approx14_pwl.txt and checked there against approx14.bin.
static const uint32_t kApproxPwl[] = {
#include "approx14_pwl.inc"
};
static_assert(sizeof(kApproxPwl) == 256 * 4, "approximation pairs");

using namespace glimp;

namespace {

// A roctx range around an entry point: the host-side zones of the
// reference's Tracy instrumentation (Rasterizer.cpp:494-612, Main.cpp:98-263),
// visible to rocprofv3 --marker-trace; inert without a tool attached.
struct Zone {
    explicit Zone(const char* name) { roctxRangePushA(name); }
    ~Zone() { roctxRangePop(); }
    Zone(const Zone&) = delete;
    Zone& operator=(const Zone&) = delete;
};

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

// Device scratch of one launch (owned by the context for single draws and
// by the batch for batches, so a captured graph keeps valid pointers).
struct Scratch {
    TriRecord* records = nullptr;
    size_t records_cap = 0;
    ClipEntry* clip = nullptr;
    size_t clip_cap = 0;
    uint2* visible_list = nullptr;
    size_t visible_list_cap = 0;
    uint32_t* meshlet_counts = nullptr;
    size_t meshlet_counts_cap = 0;
    uint32_t* bin_counts = nullptr;
    size_t bin_counts_cap = 0;
    uint32_t* bin_flags = nullptr;     // zero between draws, like bin_counts
    size_t bin_flags_cap = 0;
    uint2* clip_overflow = nullptr;
    size_t clip_overflow_cap = 0;
    uint32_t* bin_lists = nullptr;
    uint32_t* bin_order = nullptr;
    size_t bin_order_cap = 0;
    size_t bin_lists_cap = 0;
    uint32_t* spill = nullptr;         // overflow list (record indices)
    size_t spill_cap = 0;
    AlphaRecord* alpha = nullptr;      // only once a draw has alpha-tested materials
    size_t alpha_cap = 0;
    DrawScratch* scratch = nullptr;

    void release() {
        void* ptrs[] = { records, clip, visible_list, meshlet_counts, bin_counts, bin_flags, clip_overflow, bin_lists,
                         bin_order, spill, alpha, scratch };
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        *this = Scratch();
    }
};

}  // namespace

// ---------------------------------------------------------------- static scratch
//
// The scratch of glimp_draw_meshlets_static: every buffer the direct-route
// draw needs beyond its inputs and outputs (setup records and their alpha
// side records, the clip queue and its fan records, the visible list, the
// per-meshlet counts, the clip-overflow list and the draw counters), for up to
// kStaticMaxMeshlets meshlets per call. Each device holds its own copy for as
// long as the process runs. Host code never uses the array's name as an
// address: glimp_create takes this device's address with hipGetSymbolAddress
// and every use goes through that pointer.
constexpr size_t kStaticMaxMeshlets = 1024;
constexpr size_t kStaticClip = 4096;                         // clip queue entries: clip_capacity_for(1024)
constexpr size_t kStaticRecords = kStaticMaxMeshlets * 128 + kStaticClip * 8;
constexpr size_t kStaticMaxDraws = 256;                      // draws per framebuffer between clears

constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
constexpr size_t kStaticRecordsOff = 0;
constexpr size_t kStaticAlphaOff = kStaticRecordsOff + align256(kStaticRecords * sizeof(TriRecord));
constexpr size_t kStaticClipOff = kStaticAlphaOff + align256(kStaticRecords * sizeof(AlphaRecord));
constexpr size_t kStaticVisibleOff = kStaticClipOff + align256(kStaticClip * sizeof(ClipEntry));
constexpr size_t kStaticCountsOff = kStaticVisibleOff + align256(kStaticMaxMeshlets * sizeof(uint2));
constexpr size_t kStaticOverflowOff = kStaticCountsOff + align256(kStaticMaxMeshlets * sizeof(uint32_t));
constexpr size_t kStaticDrawScratchOff = kStaticOverflowOff + align256(kStaticMaxMeshlets * sizeof(uint2));
constexpr size_t kStaticBytes = kStaticDrawScratchOff + align256(sizeof(DrawScratch));   // about 16 MB

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

// this process's address of `scratch` on each device, taken once per device
std::mutex g_static_mutex;
void* g_static_ptr[64] = {};

// With `device` current: the device's address of `scratch`.
hipError_t static_scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_static_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_static_ptr[device] = ptr;
    }
    *out = g_static_ptr[device];
    return hipSuccess;
}

}  // namespace

struct glimp_context {
    int device = 0;
    uint32_t* tables = nullptr;
    Counters* counters = nullptr;
    DrawScratch* scratch = nullptr;
    uint32_t* visible = nullptr;
    uint16_t* bitmap = nullptr;        // glimp_cull_meshlets output
    size_t bitmap_cap = 0;
    Scratch draw_scratch;              // scratch of single draws
    DiffWorkspace diff_ws;             // differentiable resolve backward
    std::vector<DiffTex> diff_tex;     // last texture table uploaded to d_diff_tex
    DiffTex* d_diff_tex = nullptr;
    size_t d_diff_tex_cap = 0;
    DiffView* d_diff_views = nullptr;
    size_t d_diff_views_cap = 0;
    Profiler prof;
    // this device's copy of the static scratch array (glimp_draw_meshlets_static)
    unsigned char* static_scratch = nullptr;

    ApproxTables approx() const { return ApproxTables{ reinterpret_cast<const uint2*>(tables) }; }
};

struct glimp_texture {
    glimp_context* ctx = nullptr;
    DevTexture hdr = {};
    uint32_t* data = nullptr;
    uint32_t format = GLIMP_TEXEL_RGBA8;
};

struct glimp_scene {
    glimp_context* ctx = nullptr;
    DevMeshlets m = {};
    std::vector<void*> buffers;
    DevMaterial* materials = nullptr;
    uint32_t num_materials = 0;
    DevTexture* textures = nullptr;
    uint32_t num_textures = 0;
    bool has_alpha = false;            // a material runs the alpha test
    std::vector<uint32_t> material_id; // host copy of the meshlets' material ids
    // bad[i] = meshlets before i without a textured material (the deferred
    // pipeline samples every meshlet's texture)
    std::vector<uint32_t> untextured_before;
};

struct glimp_framebuffer {
    glimp_context* ctx = nullptr;
    uint32_t width = 0, height = 0;
    unsigned long long* keys = nullptr;
    // keys of bin-overflow fragments, zero between draws (ViewParams::spill_keys)
    unsigned long long* spill_keys = nullptr;
    uint32_t clear_color = 0;
    float clear_depth = 0.0f;
    bool pending_clear = false;        // cleared, keys not yet written
    uint32_t next_seq = 0;
    std::vector<DrawEntry> draws;
    DrawEntry* d_draws = nullptr;
    size_t d_draws_cap = 0;
    std::vector<DrawEntry> uploaded;   // contents of d_draws
    uint32_t* color = nullptr;         // resolved RGBA8, row-major
    // The pipeline whose output the framebuffer holds (GLIMP_PIPELINE_*), or
    // -1 when nothing was drawn since the clear. Overdraw keeps its counts
    // in `keys` and its depth layer in `aux`; the deferred pipeline keeps
    // its colour and normal/MR layers in `aux` (2 * width * height words).
    int content = -1;
    uint32_t* aux = nullptr;
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
    hipError_t e = upload(&c->tables, kApproxPwl, 256);
    if (e == hipSuccess) e = dev_alloc(&c->counters, 1);
    if (e == hipSuccess) e = hipMemset(c->counters, 0, sizeof(Counters));
    if (e == hipSuccess) e = dev_alloc(&c->scratch, 1);
    if (e == hipSuccess) e = hipMemset(c->scratch, 0, sizeof(DrawScratch));
    if (e == hipSuccess) e = dev_alloc(&c->visible, 1);
    void* sp = nullptr;
    if (e == hipSuccess) e = static_scratch_address(device, &sp);
    c->static_scratch = static_cast<unsigned char*>(sp);
    if (e != hipSuccess) rc = fail(GLIMP_E_HIP, std::string("glimp_create: ") + hipGetErrorString(e));
    if (rc == GLIMP_OK) warm_launch_shapes();
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
    diff_workspace_release(c->diff_ws);
    void* ptrs[] = { c->tables, c->counters, c->scratch, c->visible, c->bitmap, c->d_diff_tex, c->d_diff_views };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c->draw_scratch.release();
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
    s->material_id.assign(d->material_id, d->material_id + n);
    s->untextured_before.resize(n + 1);
    for (size_t i = 0; i <= n; i++) s->untextured_before[i] = (uint32_t)i;   // no materials yet
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
    if (d->format != GLIMP_TEXEL_RGBA8 && d->format != GLIMP_TEXEL_R11G11B10F)
        return fail(GLIMP_E_INVALID, "glimp_texture_create: unknown texel format");
    HIP_TRY(hipSetDevice(c->device));
    glimp_texture* t = new (std::nothrow) glimp_texture();
    if (!t) return fail(GLIMP_E_NOMEM, "glimp_texture_create: out of host memory");
    t->ctx = c;
    t->format = d->format;
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
        if (textures[i]->format != GLIMP_TEXEL_RGBA8)
            return fail(GLIMP_E_INVALID, "glimp_scene_set_materials: material textures are RGBA8");
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
    s->has_alpha = false;
    for (const auto& m : dm) s->has_alpha = s->has_alpha || (m.fs_id == 1 && m.texture >= 0);
    uint32_t bad = 0;
    for (size_t i = 0; i < s->material_id.size(); i++) {
        s->untextured_before[i] = bad;
        const uint32_t id = s->material_id[i];
        bad += (id >= count || dm[id].texture < 0) ? 1u : 0u;
    }
    s->untextured_before[s->material_id.size()] = bad;
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
    // room for the draw table of glimp_draw_meshlets_static, which must not
    // allocate (the table only grows past this for the other draw calls)
    f->draws.reserve(kStaticMaxDraws);
    hipError_t e = dev_alloc(&f->keys, (size_t)width * height);
    if (e == hipSuccess) e = hipMemset(f->keys, 0, (size_t)width * height * 8);
    if (e == hipSuccess) e = dev_alloc(&f->spill_keys, (size_t)width * height);
    if (e == hipSuccess) e = hipMemset(f->spill_keys, 0, (size_t)width * height * 8);
    if (e != hipSuccess) {
        if (f->keys) (void)hipFree(f->keys);
        if (f->spill_keys) (void)hipFree(f->spill_keys);
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
    if (f->spill_keys) (void)hipFree(f->spill_keys);
    if (f->d_draws) (void)hipFree(f->d_draws);
    if (f->color) (void)hipFree(f->color);
    if (f->aux) (void)hipFree(f->aux);
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
    f->content = -1;
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
// Unchanged tables are not uploaded again, so a steady clear-draw-resolve
// loop never waits for the device here.
int upload_draws(glimp_framebuffer* f, hipStream_t s, uint32_t* count) {
    std::vector<DrawEntry> draws = f->draws;
    if (draws.empty()) draws.push_back(DrawEntry{ 0, 0, 0, 0 });
    *count = (uint32_t)draws.size();
    if (f->d_draws && draws.size() == f->uploaded.size() &&
        memcmp(draws.data(), f->uploaded.data(), draws.size() * sizeof(DrawEntry)) == 0)
        return GLIMP_OK;
    f->uploaded.clear();
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
    f->uploaded = draws;
    return GLIMP_OK;
}

}  // namespace

}  // extern "C"

namespace {

// Counter words and depth of an overdraw framebuffer (host arrays, either
// may be NULL).
int read_overdraw(glimp_framebuffer* f, uint32_t* words, uint32_t* depth, int linear, hipStream_t s) {
    const size_t n = (size_t)f->width * f->height;
    uint32_t* d_out = nullptr;
    HIP_TRY(dev_alloc(&d_out, 2 * n));
    hipError_t e = launch_decode_overdraw(f->keys, f->aux, (uint32_t)n, f->clear_color, (int)f->width, d_out,
                                          d_out + n, linear, s);
    if (e == hipSuccess && words) e = hipMemcpyAsync(words, d_out, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, d_out + n, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(GLIMP_E_HIP, std::string("glimp_framebuffer_read: ") + hipGetErrorString(e));
    return GLIMP_OK;
}

// Layer `which` (0 colour, 1 normal/MR) of a deferred framebuffer.
int read_aux_layer(glimp_framebuffer* f, uint32_t which, uint32_t* out, int linear, hipStream_t s) {
    const size_t n = (size_t)f->width * f->height;
    uint32_t* d_out = nullptr;
    HIP_TRY(dev_alloc(&d_out, n));
    hipError_t e = launch_detile32(f->aux + which * n, (int)f->width, (uint32_t)n, d_out, linear, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(GLIMP_E_HIP, std::string("glimp_framebuffer_read: ") + hipGetErrorString(e));
    return GLIMP_OK;
}

}  // namespace

extern "C" {

int glimp_framebuffer_read(glimp_framebuffer* f, uint32_t* ids, float* depth, int linear, void* stream) {
    if (!f) return fail(GLIMP_E_INVALID, "glimp_framebuffer_read: NULL framebuffer");
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    size_t n = (size_t)f->width * f->height;
    if (f->content == (int)GLIMP_PIPELINE_OVERDRAW) return read_overdraw(f, ids, (uint32_t*)depth, linear, s);
    if (f->content == (int)GLIMP_PIPELINE_DEFERRED && ids) {
        // layer 0 is the G-buffer's colour layer; depth still comes from the keys
        int rc = read_aux_layer(f, 0, ids, linear, s);
        if (rc != GLIMP_OK || !depth) return rc;
        ids = nullptr;
    }
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

int glimp_framebuffer_read_layer(glimp_framebuffer* f, uint32_t layer, uint32_t* out, int linear, void* stream) {
    if (!f || !out) return fail(GLIMP_E_INVALID, "glimp_framebuffer_read_layer: NULL argument");
    if (layer == 0) return glimp_framebuffer_read(f, out, nullptr, linear, stream);
    if (layer == 1) return glimp_framebuffer_read(f, nullptr, (float*)out, linear, stream);
    if (layer == 2 && f->aux && f->content != (int)GLIMP_PIPELINE_OVERDRAW) {
        HIP_TRY(hipSetDevice(f->ctx->device));
        return read_aux_layer(f, 1, out, linear, (hipStream_t)stream);
    }
    return fail(GLIMP_E_INVALID, "glimp_framebuffer_read_layer: the framebuffer has no such layer");
}

}  // extern "C"

struct glimp_hiz {
    glimp_context* ctx = nullptr;
    Occlusion layout = {};             // depth map pointer, sizes and level table
    float* data = nullptr;
    size_t floats = 0;
};

namespace {

// Depth-map layout for a width x height frame: the reference's level count
// (power-of-two sides above (size - 1) / 2, levels while both sides are at
// least 4, Main.cpp:54-56 and Texture.h:606-611) and levels 0 .. count - 2
// stored row-major.
size_t hiz_layout(uint32_t W, uint32_t H, Occlusion& o) {
    auto half = [](uint32_t v) { uint32_t h = (v - 1) / 2; return h ? 1u << (32 - __builtin_clz(h)) : 1u; };
    const uint32_t hw = half(W), hh = half(H);
    uint32_t levels = 0;
    for (; levels < 16; levels++)
        if ((hw >> levels) < 4 || (hh >> levels) < 4) break;
    o.mip_levels = levels;
    o.width = W;
    o.height = H;
    size_t off = 0;
    for (uint32_t m = 0; m + 1 < levels; m++) {
        o.level_offset[m] = (uint32_t)off;
        o.level_width[m] = ((W - 1) >> (m + 1)) + 1;
        off += (size_t)o.level_width[m] * (((H - 1) >> (m + 1)) + 1);
    }
    return off;
}

int make_occlusion(const glimp_occlusion* d, const std::string& who, Occlusion& o) {
    if (!d->depth_map) return fail(GLIMP_E_INVALID, who + ": occlusion without a depth map");
    o = d->depth_map->layout;
    memcpy(o.m, d->object_to_prev_view, sizeof(o.m));
    o.scale = d->object_to_world_scale;
    memcpy(o.proj, d->proj_factors, sizeof(o.proj));
    return GLIMP_OK;
}

}  // namespace

namespace {

struct LaunchShape {
    uint32_t bin_size = 16;
    uint32_t bin_capacity = 1024;
    uint32_t raster_threads = 256;
};

// Clipped triangles are rare (the reference measures < 0.1%, README.md:93);
// the clip queue holds up to two per meshlet and every queued triangle may
// produce 8 fan triangles. What a full queue drops is counted in the
// context's `dropped` counter (glimp_get_counters).
size_t clip_capacity_for(size_t meshlets) { return std::max<size_t>(4096, meshlets * 2); }

// A batch's views share one queue: as large as theirs drawn one at a time
// would be together, up to 2^20 triangles (64 MiB of queue, 320 MiB of fan
// records), and never smaller than the single-draw rule for all its meshlets.
constexpr size_t kBatchClipCap = size_t(1) << 20;
size_t batch_clip_capacity(const std::vector<glimp_view_desc>& views, size_t total_meshlets) {
    size_t sum = 0;
    for (const auto& d : views) sum = std::min(sum + clip_capacity_for(d.meshlet_count), kBatchClipCap);
    return std::max(clip_capacity_for(total_meshlets), sum);
}

// One view of a draw (ShadingContext + DrawMeshlets arguments).
void make_view(glimp_framebuffer* f, glimp_scene* s, uint32_t meshlet_offset, uint32_t meshlet_count,
               const float* clip, const float* planes, uint32_t flags, ViewParams& V, const Occlusion* occ = nullptr) {
    memset(&V, 0, sizeof(V));
    V.m = s->m;
    V.materials = s->materials;
    V.num_materials = s->num_materials;
    V.textures = s->textures;
    V.meshlet_offset = meshlet_offset;
    V.meshlet_count = meshlet_count;
    V.seq_base = f->next_seq;
    V.width = (int)f->width;
    V.height = (int)f->height;
    bool guard = !(flags & GLIMP_DRAW_NO_GUARDBAND);
    V.guard_x = guard ? (float)GLIMP_MAX_RENDER_SIZE / (float)f->width : 1.0f;
    V.guard_y = guard ? (float)GLIMP_MAX_RENDER_SIZE / (float)f->height : 1.0f;
    V.clipping = !(flags & GLIMP_DRAW_NO_CLIPPING);
    V.use_cull = planes != nullptr;
    if (planes) memcpy(V.planes, planes, sizeof(V.planes));
    if (planes && occ) V.occ = *occ;
    memcpy(V.clip, clip, sizeof(V.clip));
    V.clear_depth = f->clear_depth;
    V.fresh = f->pending_clear ? 1 : 0;
    V.keys = f->keys;
    V.spill_keys = f->spill_keys;
}

// Lays out bins for `views` (bin_base, bins_x/y) and returns the total.
uint32_t layout_bins(std::vector<ViewParams>& views, uint32_t bin_size) {
    uint32_t base = 0;
    for (auto& V : views) {
        V.bin_base = base;
        V.bins_x = ((uint32_t)V.width + bin_size - 1) / bin_size;
        V.bins_y = ((uint32_t)V.height + bin_size - 1) / bin_size;
        base += V.bins_x * V.bins_y;
    }
    return base;
}

// Bin indices are packed with a 6-bit lane number into 32 bits when binning.
constexpr uint64_t kMaxBins = 1ull << 26;

uint64_t count_bins(const std::vector<ViewParams>& views, uint32_t bin_size) {
    uint64_t n = 0;
    for (const auto& V : views)
        n += (uint64_t)(((uint32_t)V.width + bin_size - 1) / bin_size) * (((uint32_t)V.height + bin_size - 1) / bin_size);
    return n;
}

// The overflow list takes each record at most once: one entry per record slot.
size_t spill_capacity_for(size_t record_slots) { return record_slots; }

int prepare_scratch(Scratch& S, size_t total_meshlets, size_t clipq, size_t total_bins, const LaunchShape& sh,
                    bool binned, bool alpha) {
    const size_t recs = total_meshlets * 128 + clipq * 8;
    int rc = grow_one(S.records, S.records_cap, recs);
    if (rc == GLIMP_OK) rc = grow_one(S.clip, S.clip_cap, clipq);
    if (rc == GLIMP_OK) rc = grow_one(S.visible_list, S.visible_list_cap, std::max<size_t>(total_meshlets, 1));
    if (rc == GLIMP_OK) rc = grow_one(S.meshlet_counts, S.meshlet_counts_cap, total_meshlets);
    if (rc == GLIMP_OK) rc = grow_one(S.clip_overflow, S.clip_overflow_cap, std::max<size_t>(total_meshlets, 1));
    const size_t old_counts_cap = S.bin_counts_cap;
    if (rc == GLIMP_OK) rc = grow_one(S.bin_counts, S.bin_counts_cap, std::max<size_t>(total_bins, 1));
    // draws leave the bin counts cleared; a regrown buffer starts cleared (the
    // test is on the capacity: hipMalloc may hand back the freed address)
    if (rc == GLIMP_OK && S.bin_counts_cap != old_counts_cap)
        HIP_TRY(hipMemset(S.bin_counts, 0, S.bin_counts_cap * sizeof(uint32_t)));
    const size_t old_flags_cap = S.bin_flags_cap;
    if (rc == GLIMP_OK && binned) rc = grow_one(S.bin_flags, S.bin_flags_cap, std::max<size_t>(total_bins, 1));
    if (rc == GLIMP_OK && S.bin_flags_cap != old_flags_cap)
        HIP_TRY(hipMemset(S.bin_flags, 0, S.bin_flags_cap * sizeof(uint32_t)));
    if (rc == GLIMP_OK && binned) rc = grow_one(S.bin_lists, S.bin_lists_cap, total_bins * sh.bin_capacity);
    if (rc == GLIMP_OK && binned) rc = grow_one(S.bin_order, S.bin_order_cap, total_bins * kBinBuckets);
    if (rc == GLIMP_OK && binned) rc = grow_one(S.spill, S.spill_cap, spill_capacity_for(recs));
    if (rc == GLIMP_OK && alpha) rc = grow_one(S.alpha, S.alpha_cap, S.records_cap);
    if (rc == GLIMP_OK && !S.scratch) {
        HIP_TRY(dev_alloc(&S.scratch, 1));
        HIP_TRY(hipMemset(S.scratch, 0, sizeof(DrawScratch)));
    }
    return rc;
}

void fill_params(glimp_context* c, const Scratch& S, size_t total_meshlets, size_t clipq, uint32_t total_bins,
                 const LaunchShape& sh, uint32_t flags, bool alpha, DrawParams& p) {
    p.total_meshlets = (uint32_t)total_meshlets;
    p.total_bins = total_bins;
    p.flags = flags;
    p.tables = c->approx();
    p.visible_list = S.visible_list;
    p.records = S.records;
    p.alpha = alpha ? S.alpha : nullptr;
    p.meshlet_counts = S.meshlet_counts;
    p.clip_record_base = (uint32_t)(total_meshlets * 128);
    // the draw's own capacities (not what earlier, larger draws allocated), so
    // what a full queue drops does not depend on the context's history
    p.clip_record_capacity =
        (uint32_t)std::min<size_t>(std::min(clipq * 8, S.records_cap - total_meshlets * 128), 0x7FFFFFFFu);
    p.clip_queue = S.clip;
    p.clip_capacity = (uint32_t)std::min<size_t>(std::min(clipq, S.clip_cap), 0x7FFFFFFFu);
    p.scratch = S.scratch;
    p.counters = c->counters;
    p.bin_size = sh.bin_size;
    p.bin_shift = (uint32_t)__builtin_ctz(sh.bin_size);
    p.bin_capacity = sh.bin_capacity;
    p.raster_threads = sh.raster_threads;
    p.bin_counts = S.bin_counts;
    p.bin_lists = S.bin_lists;
    p.bin_order = S.bin_order;
    p.spill = S.spill;
    p.spill_capacity = (uint32_t)std::min(spill_capacity_for(total_meshlets * 128 + clipq * 8), S.spill_cap);
    p.bin_flags = S.bin_flags;
    p.clip_overflow = S.clip_overflow;
    p.exact_overflow = 1u;          // vis-buffer pipeline; the overdraw and deferred draws turn it off
}

bool valid_bin_size(uint32_t b) { return b == 8 || b == 16 || b == 32 || b == 64; }

}  // namespace

extern "C" {

int glimp_draw_meshlets(glimp_context* c, glimp_framebuffer* f, glimp_scene* s, const glimp_draw_desc* d,
                        void* stream) {
    Zone zone("glimp_draw_meshlets");
    if (!c || !f || !s || !d) return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: NULL argument");
    if (d->pipeline != GLIMP_PIPELINE_VISBUFFER && d->pipeline != GLIMP_PIPELINE_OVERDRAW &&
        d->pipeline != GLIMP_PIPELINE_DEFERRED)
        return fail(GLIMP_E_UNSUPPORTED, "glimp_draw_meshlets: unknown pipeline");
    if (f->content >= 0 && f->content != (int)d->pipeline)
        return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: the framebuffer holds another pipeline's output; clear it first");
    if ((uint64_t)d->meshlet_offset + d->meshlet_count > s->m.count)
        return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: meshlet range outside the scene");
    if ((uint64_t)f->next_seq + d->meshlet_count > GLIMP_MAX_DRAWN_MESHLETS)
        return fail(GLIMP_E_LIMIT, "glimp_draw_meshlets: more than GLIMP_MAX_DRAWN_MESHLETS meshlets since the last clear");
    LaunchShape sh;
    if (d->bin_size) sh.bin_size = d->bin_size;
    if (d->bin_capacity) sh.bin_capacity = d->bin_capacity;
    if (!valid_bin_size(sh.bin_size))
        return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: bin_size must be 8, 16, 32 or 64");
    if (d->meshlet_count == 0) return GLIMP_OK;
    HIP_TRY(hipSetDevice(c->device));
    const bool overdraw = d->pipeline == GLIMP_PIPELINE_OVERDRAW;
    const bool deferred = d->pipeline == GLIMP_PIPELINE_DEFERRED;
    const bool direct = overdraw || deferred || (d->flags & GLIMP_DRAW_DIRECT) != 0;
    hipStream_t st = (hipStream_t)stream;
    if (deferred) {
        // FS_EncodeGBuffer samples the texture of every meshlet's material
        // (Shading.cpp:357-366); without one the reference reads through a
        // null texture, and a meshlet without a material leaves the colour of
        // whichever earlier fragment passed the depth test, which one pass
        // per draw cannot reproduce: such ranges are refused
        if (s->untextured_before[d->meshlet_offset + d->meshlet_count] != s->untextured_before[d->meshlet_offset])
            return fail(GLIMP_E_UNSUPPORTED,
                        "glimp_draw_meshlets: the deferred pipeline needs a textured material on every meshlet");
        const size_t n = (size_t)f->width * f->height;
        if (!f->aux) {
            HIP_TRY(dev_alloc(&f->aux, 2 * n));
            HIP_TRY(hipMemsetAsync(f->aux, 0, 2 * n * 4, st));
        }
        if (f->content != (int)GLIMP_PIPELINE_DEFERRED)   // Clear's colour; layer 2 keeps its contents
            HIP_TRY(launch_fill32(f->aux, (uint32_t)n, f->clear_color, st));
    }
    if (overdraw && f->content != (int)GLIMP_PIPELINE_OVERDRAW) {
        // first overdraw draw since the clear: zero counts, depth = clear depth
        const size_t n = (size_t)f->width * f->height;
        if (!f->aux) HIP_TRY(dev_alloc(&f->aux, 2 * n));
        uint32_t db;
        memcpy(&db, &f->clear_depth, 4);
        HIP_TRY(launch_clear(f->keys, (uint32_t)n, 0ull, st));
        HIP_TRY(launch_fill32(f->aux, (uint32_t)n, db, st));
        f->pending_clear = false;
    }
    if (direct) HIP_TRY(flush_clear(f, st));

    Occlusion occ;
    if (d->occlusion) {
        if (!d->frustum_planes) return fail(GLIMP_E_INVALID, "glimp_draw_meshlets: occlusion needs frustum planes");
        int orc = make_occlusion(d->occlusion, "glimp_draw_meshlets", occ);
        if (orc != GLIMP_OK) return orc;
    }
    std::vector<ViewParams> views(1);
    make_view(f, s, d->meshlet_offset, d->meshlet_count, d->object_to_clip, d->frustum_planes, d->flags, views[0],
              d->occlusion ? &occ : nullptr);
    const uint32_t total_bins = direct ? 0 : layout_bins(views, sh.bin_size);
    const bool alpha = (s->has_alpha && !overdraw) || deferred;    // FS_Overdraw has no alpha test
    const size_t clipq = clip_capacity_for(d->meshlet_count);
    int rc = prepare_scratch(c->draw_scratch, d->meshlet_count, clipq, total_bins, sh, !direct, alpha);
    if (rc != GLIMP_OK) return rc;
    DrawParams p;
    memset(&p, 0, sizeof(p));
    fill_params(c, c->draw_scratch, d->meshlet_count, clipq, total_bins, sh, d->flags, alpha, p);
    p.num_views = 1;
    p.views = nullptr;
    p.view = views[0];
    p.view.aux = f->aux;
    memcpy(p.view.o2w, d->object_to_world, sizeof(p.view.o2w));
    p.deferred = deferred ? 1u : 0u;
    // clipped triangles without queue or record room are drawn at once only by
    // the vis-buffer pipeline; the others count them in `dropped`
    p.exact_overflow = (overdraw || deferred) ? 0u : 1u;
    if (overdraw) HIP_TRY(launch_draw_overdraw(p, st, &c->prof));
    else if (deferred) HIP_TRY(launch_draw_deferred(p, st, &c->prof));
    else if (direct) HIP_TRY(launch_draw_direct(p, st, &c->prof));
    else HIP_TRY(launch_draw_binned(p, st, &c->prof));
    f->pending_clear = false;
    f->content = (int)d->pipeline;
    f->draws.push_back(DrawEntry{ f->next_seq, d->meshlet_count, d->meshlet_offset, 0 });
    f->next_seq += d->meshlet_count;
    return GLIMP_OK;
}

int glimp_draw_meshlets_static(glimp_context* c, glimp_framebuffer* f, glimp_scene* s, const glimp_draw_desc* d,
                               void* stream) {
    if (!c || !f || !s || !d) return fail(GLIMP_E_INVALID, "glimp_draw_meshlets_static: NULL argument");
    if (d->pipeline != GLIMP_PIPELINE_VISBUFFER)
        return fail(GLIMP_E_UNSUPPORTED, "glimp_draw_meshlets_static: only the vis-buffer pipeline");
    if (f->content >= 0 && f->content != (int)GLIMP_PIPELINE_VISBUFFER)
        return fail(GLIMP_E_INVALID,
                    "glimp_draw_meshlets_static: the framebuffer holds another pipeline's output; clear it first");
    if ((uint64_t)d->meshlet_offset + d->meshlet_count > s->m.count)
        return fail(GLIMP_E_INVALID, "glimp_draw_meshlets_static: meshlet range outside the scene");
    if (d->meshlet_count > kStaticMaxMeshlets)
        return fail(GLIMP_E_LIMIT, "glimp_draw_meshlets_static: more than 1024 meshlets in one call "
                                   "(the size of the static scratch); split the range or use glimp_draw_meshlets");
    if ((uint64_t)f->next_seq + d->meshlet_count > GLIMP_MAX_DRAWN_MESHLETS)
        return fail(GLIMP_E_LIMIT,
                    "glimp_draw_meshlets_static: more than GLIMP_MAX_DRAWN_MESHLETS meshlets since the last clear");
    if (f->draws.size() >= f->draws.capacity())
        return fail(GLIMP_E_LIMIT, "glimp_draw_meshlets_static: more than 256 draws since the last clear");
    if (!c->static_scratch) return fail(GLIMP_E_HIP, "glimp_draw_meshlets_static: no static scratch on this device");
    if (d->meshlet_count == 0) return GLIMP_OK;
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(flush_clear(f, st));
    Occlusion occ;
    if (d->occlusion) {
        if (!d->frustum_planes) return fail(GLIMP_E_INVALID, "glimp_draw_meshlets_static: occlusion needs frustum planes");
        int orc = make_occlusion(d->occlusion, "glimp_draw_meshlets_static", occ);
        if (orc != GLIMP_OK) return orc;
    }
    unsigned char* base = c->static_scratch;
    const size_t count = d->meshlet_count;
    const size_t clipq = clip_capacity_for(count);
    DrawParams p;
    memset(&p, 0, sizeof(p));
    make_view(f, s, d->meshlet_offset, d->meshlet_count, d->object_to_clip, d->frustum_planes, d->flags, p.view,
              d->occlusion ? &occ : nullptr);
    p.num_views = 1;
    p.views = nullptr;
    p.total_meshlets = (uint32_t)count;
    p.total_bins = 0;                                  // the direct route
    p.flags = d->flags | GLIMP_DRAW_DIRECT;
    p.tables = c->approx();
    p.records = reinterpret_cast<TriRecord*>(base + kStaticRecordsOff);
    p.alpha = s->has_alpha ? reinterpret_cast<AlphaRecord*>(base + kStaticAlphaOff) : nullptr;
    p.clip_queue = reinterpret_cast<ClipEntry*>(base + kStaticClipOff);
    p.visible_list = reinterpret_cast<uint2*>(base + kStaticVisibleOff);
    p.meshlet_counts = reinterpret_cast<uint32_t*>(base + kStaticCountsOff);
    p.clip_overflow = reinterpret_cast<uint2*>(base + kStaticOverflowOff);
    p.scratch = reinterpret_cast<DrawScratch*>(base + kStaticDrawScratchOff);
    p.clip_record_base = (uint32_t)(count * 128);
    p.clip_record_capacity = (uint32_t)(clipq * 8);
    p.clip_capacity = (uint32_t)clipq;
    p.counters = c->counters;
    p.exact_overflow = 1u;
    // the array keeps the last call's contents: the counters start at zero
    // (a kernel, not a memset, so the call can be captured)
    HIP_TRY(launch_reset_counters(p.scratch, st));
    HIP_TRY(launch_draw_direct(p, st, nullptr));
    f->pending_clear = false;
    f->content = (int)GLIMP_PIPELINE_VISBUFFER;
    f->draws.push_back(DrawEntry{ f->next_seq, d->meshlet_count, d->meshlet_offset, 0 });   // capacity reserved
    f->next_seq += d->meshlet_count;
    return GLIMP_OK;
}

}  // extern "C"

struct glimp_batch {
    glimp_context* ctx = nullptr;
    std::vector<glimp_view_desc> descs;
    std::vector<glimp_occlusion> occlusion;   // per view, copied (descs point here)
    std::vector<ViewParams> views;
    ViewParams* d_views = nullptr;
    size_t total_meshlets = 0;
    size_t clip_triangles = 0;         // clip queue capacity (batch_clip_capacity)
    uint32_t total_bins = 0;
    LaunchShape shape;
    bool alpha = false;                // a view's scene has alpha-tested materials
    Scratch scratch;
    hipStream_t stream = nullptr;      // capture stream (replays run on the caller's stream)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // resolves after the draw (glimp_batch_set_resolve), with copies of
    // their light arrays and the batch's own one-entry draw tables
    std::vector<glimp_resolve_desc> resolves;
    std::vector<std::vector<glimp_light>> resolve_lights;
    DrawEntry* d_tables = nullptr;
};

namespace {

int build_resolve(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_resolve_desc* d, hipStream_t s,
                  ResolveParams& p, const std::string& who);

void drop_graph(glimp_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Rebuilds the view table for the batch's shape and uploads it.
int layout_batch(glimp_batch* b) {
    std::vector<ViewParams> probe = b->views;
    if (count_bins(probe, b->shape.bin_size) >= kMaxBins)
        return fail(GLIMP_E_LIMIT, "glimp_batch: more than 2^26 screen bins; use a larger bin size or fewer views");
    b->total_bins = layout_bins(b->views, b->shape.bin_size);
    b->alpha = false;
    for (const auto& d : b->descs) b->alpha = b->alpha || d.scene->has_alpha;
    int rc = prepare_scratch(b->scratch, b->total_meshlets, b->clip_triangles, b->total_bins, b->shape, true,
                             b->alpha);
    if (rc != GLIMP_OK) return rc;
    if (!b->d_views) HIP_TRY(dev_alloc(&b->d_views, b->views.size()));
    HIP_TRY(hipMemcpy(b->d_views, b->views.data(), b->views.size() * sizeof(ViewParams), hipMemcpyHostToDevice));
    drop_graph(b);
    return GLIMP_OK;
}

void batch_params(glimp_batch* b, DrawParams& p) {
    memset(&p, 0, sizeof(p));
    fill_params(b->ctx, b->scratch, b->total_meshlets, b->clip_triangles, b->total_bins, b->shape, 0, b->alpha, p);
    p.num_views = (uint32_t)b->views.size();
    p.views = b->d_views;
}

bool valid_shape(const glimp_batch_shape& s) {
    const bool wide = s.raster_threads == 512 || s.raster_threads == 1024;    // 32- and 64-pixel bins only
    return valid_bin_size(s.bin_size) && s.bin_capacity >= 1 &&
           (s.raster_threads == 128 || s.raster_threads == 256 || (wide && s.bin_size >= 32));
}

// Framebuffer state after a batch launch: cleared, then one draw.
void mark_drawn(glimp_batch* b) {
    for (const auto& d : b->descs) {
        glimp_framebuffer* f = d.framebuffer;
        f->clear_color = d.clear_color;
        f->clear_depth = d.clear_depth == 0.0f ? 0.0f : d.clear_depth;
        f->pending_clear = false;
        f->content = (int)GLIMP_PIPELINE_VISBUFFER;
        f->draws.assign(1, DrawEntry{ 0, d.meshlet_count, d.meshlet_offset, 0 });
        f->next_seq = d.meshlet_count;
    }
}

}  // namespace

extern "C" {

int glimp_batch_create(glimp_context* c, const glimp_view_desc* views, uint32_t count, const glimp_batch_shape* shape,
                       glimp_batch** out) {
    if (!c || !out || (count && !views)) return fail(GLIMP_E_INVALID, "glimp_batch_create: NULL argument");
    *out = nullptr;
    if (count == 0) return fail(GLIMP_E_INVALID, "glimp_batch_create: no views");
    if (count > 65535) return fail(GLIMP_E_LIMIT, "glimp_batch_create: at most 65535 views");
    std::vector<glimp_framebuffer*> seen;
    for (uint32_t i = 0; i < count; i++) {
        const glimp_view_desc& d = views[i];
        if (!d.scene || !d.framebuffer) return fail(GLIMP_E_INVALID, "glimp_batch_create: view without scene or framebuffer");
        if ((uint64_t)d.meshlet_offset + d.meshlet_count > d.scene->m.count)
            return fail(GLIMP_E_INVALID, "glimp_batch_create: meshlet range outside the scene");
        if (d.meshlet_count > GLIMP_MAX_DRAWN_MESHLETS) return fail(GLIMP_E_LIMIT, "glimp_batch_create: view too large");
        if (!(d.clear_depth >= 0.0f)) return fail(GLIMP_E_INVALID, "glimp_batch_create: clear depth must be >= 0");
        if (d.occlusion && !d.frustum_planes)
            return fail(GLIMP_E_INVALID, "glimp_batch_create: occlusion needs frustum planes");
        if (d.occlusion && !d.occlusion->depth_map)
            return fail(GLIMP_E_INVALID, "glimp_batch_create: occlusion without a depth map");
        for (auto* f : seen)
            if (f == d.framebuffer) return fail(GLIMP_E_INVALID, "glimp_batch_create: views must not share a framebuffer");
        seen.push_back(d.framebuffer);
    }
    HIP_TRY(hipSetDevice(c->device));
    glimp_batch* b = new (std::nothrow) glimp_batch();
    if (!b) return fail(GLIMP_E_NOMEM, "glimp_batch_create: out of host memory");
    b->ctx = c;
    if (shape) {
        if (!valid_shape(*shape)) {
            delete b;
            return fail(GLIMP_E_INVALID, "glimp_batch_create: bad shape");
        }
        b->shape.bin_size = shape->bin_size;
        b->shape.bin_capacity = shape->bin_capacity;
        b->shape.raster_threads = shape->raster_threads;
    }
    b->descs.assign(views, views + count);
    b->occlusion.resize(count);
    for (uint32_t i = 0; i < count; i++) {
        if (!views[i].occlusion) continue;
        b->occlusion[i] = *views[i].occlusion;
        b->descs[i].occlusion = &b->occlusion[i];
    }
    b->views.resize(count);
    size_t base = 0;
    for (uint32_t i = 0; i < count; i++) {
        const glimp_view_desc& d = b->descs[i];
        Occlusion occ;
        if (d.occlusion) (void)make_occlusion(d.occlusion, "glimp_batch_create", occ);
        glimp_framebuffer tmp;     // the view starts from a cleared framebuffer
        tmp.width = d.framebuffer->width;
        tmp.height = d.framebuffer->height;
        tmp.keys = d.framebuffer->keys;
        tmp.spill_keys = d.framebuffer->spill_keys;
        tmp.clear_depth = d.clear_depth == 0.0f ? 0.0f : d.clear_depth;
        tmp.pending_clear = true;
        tmp.next_seq = 0;
        make_view(&tmp, d.scene, d.meshlet_offset, d.meshlet_count, d.object_to_clip, d.frustum_planes, d.flags,
                  b->views[i], d.occlusion ? &occ : nullptr);
        b->views[i].meshlet_base = (uint32_t)base;
        base += d.meshlet_count;
    }
    b->total_meshlets = base;
    b->clip_triangles = batch_clip_capacity(b->descs, base);
    // record indices (meshlet slots of 128, then 8 fan records per queued
    // triangle) are 32-bit and must stay below 2^31
    if (b->total_meshlets * 128 + b->clip_triangles * 8 >= (size_t(1) << 31)) {
        glimp_batch_destroy(b);
        return fail(GLIMP_E_LIMIT, "glimp_batch_create: the views hold too many meshlets together "
                                   "(meshlets * 128 + clip queue * 8 must stay below 2^31)");
    }
    int rc = layout_batch(b);
    hipError_t e = hipSuccess;
    if (rc == GLIMP_OK) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (rc == GLIMP_OK && e != hipSuccess) rc = fail(GLIMP_E_HIP, std::string("glimp_batch_create: ") + hipGetErrorString(e));
    if (rc != GLIMP_OK) {
        glimp_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GLIMP_OK;
}

void glimp_batch_destroy(glimp_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    // replays run on the callers' streams: wait for the device before the
    // graph and the buffers it uses go away
    (void)hipDeviceSynchronize();
    drop_graph(b);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_tables) (void)hipFree(b->d_tables);
    b->scratch.release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int glimp_batch_set_shape(glimp_batch* b, const glimp_batch_shape* shape) {
    if (!b || !shape) return fail(GLIMP_E_INVALID, "glimp_batch_set_shape: NULL argument");
    if (!valid_shape(*shape)) return fail(GLIMP_E_INVALID, "glimp_batch_set_shape: bad shape");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipDeviceSynchronize());   // scratch may be reallocated under launches on any stream
    b->shape.bin_size = shape->bin_size;
    b->shape.bin_capacity = shape->bin_capacity;
    b->shape.raster_threads = shape->raster_threads;
    return layout_batch(b);
}

int glimp_batch_launch(glimp_batch* b, int use_graph, void* stream) {
    Zone zone("glimp_batch_launch");
    if (!b) return fail(GLIMP_E_INVALID, "glimp_batch_launch: NULL batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t user = (hipStream_t)stream;
    DrawParams p;
    batch_params(b, p);
    if (!use_graph) {
        HIP_TRY(launch_draw_binned(p, user, &b->ctx->prof));
        mark_drawn(b);
        for (size_t v = 0; v < b->resolves.size(); v++) {
            int rc = glimp_resolve(b->ctx, b->descs[v].framebuffer, b->descs[v].scene, &b->resolves[v], stream);
            if (rc != GLIMP_OK) return rc;
        }
        return GLIMP_OK;
    }
    if (!b->exec) {
        // The captured resolves read each view as the batch leaves it: its
        // keys, its clear depth and a one-entry draw table owned by the batch
        // (the framebuffer's own table may change between replays).
        std::vector<ResolveParams> rps(b->resolves.size());
        if (!b->resolves.empty()) {
            mark_drawn(b);
            if (!b->d_tables) {
                std::vector<DrawEntry> t(b->descs.size());
                for (size_t v = 0; v < t.size(); v++)
                    t[v] = DrawEntry{ 0, b->descs[v].meshlet_count, b->descs[v].meshlet_offset, 0 };
                HIP_TRY(upload(&b->d_tables, t.data(), t.size()));
            }
            for (size_t v = 0; v < rps.size(); v++) {
                const glimp_view_desc& d = b->descs[v];
                // on the caller's stream: after work in flight that reads the framebuffer
                int rc = build_resolve(b->ctx, d.framebuffer, d.scene, &b->resolves[v], user, rps[v],
                                       "glimp_batch_launch");
                if (rc != GLIMP_OK) return rc;
                rps[v].draws = b->d_tables + v;
                rps[v].num_draws = 1;
            }
        }
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_draw_binned(p, b->stream, nullptr);
        for (size_t v = 0; v < rps.size() && le == hipSuccess; v++) le = launch_resolve(rps[v], b->stream, nullptr);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(GLIMP_E_HIP, std::string("glimp_batch_launch: capture failed: ") +
                                         hipGetErrorString(le != hipSuccess ? le : ce));
        }
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    Profiler::Scope sc(&b->ctx->prof, "batch_graph", user);
    HIP_TRY(hipGraphLaunch(b->exec, user));
    mark_drawn(b);
    return GLIMP_OK;
}

int glimp_batch_set_resolve(glimp_batch* b, const glimp_resolve_desc* descs) {
    if (!b) return fail(GLIMP_E_INVALID, "glimp_batch_set_resolve: NULL batch");
    std::vector<glimp_resolve_desc> r;
    std::vector<std::vector<glimp_light>> lights;
    if (descs) {
        r.assign(descs, descs + b->descs.size());
        lights.resize(r.size());
        for (size_t v = 0; v < r.size(); v++) {
            if (r[v].mode > GLIMP_RESOLVE_TRIANGLE_ID)
                return fail(GLIMP_E_INVALID, "glimp_batch_set_resolve: unknown mode");
            if (r[v].num_lights > GLIMP_MAX_LIGHTS || (r[v].num_lights && !r[v].lights))
                return fail(GLIMP_E_LIMIT, "glimp_batch_set_resolve: at most GLIMP_MAX_LIGHTS lights");
            lights[v].assign(r[v].lights, r[v].lights + r[v].num_lights);
            r[v].lights = lights[v].data();
        }
    }
    b->resolves.swap(r);
    b->resolve_lights.swap(lights);
    drop_graph(b);
    return GLIMP_OK;
}

int glimp_batch_launch_per_view(glimp_batch* b, void* stream) {
    if (!b) return fail(GLIMP_E_INVALID, "glimp_batch_launch_per_view: NULL batch");
    for (const auto& d : b->descs) {
        int rc = glimp_framebuffer_clear(d.framebuffer, d.clear_color, d.clear_depth, stream);
        if (rc != GLIMP_OK) return rc;
        glimp_draw_desc dd;
        memset(&dd, 0, sizeof(dd));
        dd.meshlet_offset = d.meshlet_offset;
        dd.meshlet_count = d.meshlet_count;
        memcpy(dd.object_to_clip, d.object_to_clip, sizeof(dd.object_to_clip));
        dd.frustum_planes = d.frustum_planes;
        dd.flags = d.flags & (GLIMP_DRAW_NO_CLIPPING | GLIMP_DRAW_NO_GUARDBAND);
        dd.pipeline = GLIMP_PIPELINE_VISBUFFER;
        dd.occlusion = d.occlusion;
        dd.bin_size = b->shape.bin_size;
        dd.bin_capacity = b->shape.bin_capacity;
        rc = glimp_draw_meshlets(b->ctx, d.framebuffer, d.scene, &dd, stream);
        if (rc != GLIMP_OK) return rc;
    }
    for (size_t v = 0; v < b->resolves.size(); v++) {
        int rc = glimp_resolve(b->ctx, b->descs[v].framebuffer, b->descs[v].scene, &b->resolves[v], stream);
        if (rc != GLIMP_OK) return rc;
    }
    return GLIMP_OK;
}

int glimp_batch_autotune(glimp_batch* b, void* stream, glimp_batch_shape* chosen, float* best_ms) {
    if (!b) return fail(GLIMP_E_INVALID, "glimp_batch_autotune: NULL batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // bin sizes in the order they usually win; capacities from the largest
    // down, so the first launch of a bin size tells the largest bin's record
    // count: smaller capacities that would overflow are not launched, and
    // when even the largest overflows only it is tried (the overflowing
    // records are drawn by the overflow raster, at a measured cost)
    const uint32_t sizes[] = { 16, 32, 8, 64 };
    const uint32_t caps[] = { 4096, 2048, 1024, 512, 256 };
    const uint32_t threads[] = { 128, 256, 512, 1024 };
    const size_t kListBudget = (size_t)2 << 30;   // bytes of bin lists a candidate may use
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    glimp_batch_shape best = { b->shape.bin_size, b->shape.bin_capacity, b->shape.raster_threads };
    float bestMs = 1e30f;
    bool found = false;
    const bool prof = b->ctx->prof.enabled;
    b->ctx->prof.enabled = false;
    int rc = GLIMP_OK;
    auto timed = [&](int reps, float& ms) -> int {
        (void)hipEventRecord(e0, st);
        int r = GLIMP_OK;
        for (int i = 0; i < reps && r == GLIMP_OK; i++) r = glimp_batch_launch(b, 0, stream);
        (void)hipEventRecord(e1, st);
        if (r != GLIMP_OK) return r;
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        return GLIMP_OK;
    };
    for (uint32_t bs : sizes) {
        uint32_t maxCount = 0;            // the largest bin at this bin size, once known
        bool firstOfSize = true;
        bool skipSize = false;
        for (uint32_t cap : caps) {
            if (skipSize || (maxCount && cap < maxCount && cap < caps[0])) continue;
            for (uint32_t th : threads) {
                const uint64_t bins = count_bins(b->views, bs);
                if (bins >= kMaxBins || bins * cap * 4 > kListBudget) continue;
                glimp_batch_shape sh = { bs, cap, th };
                if (!valid_shape(sh)) continue;
                rc = glimp_batch_set_shape(b, &sh);
                if (rc != GLIMP_OK) break;
                // One probe launch: a shape that drops anything is rejected,
                // one slower than twice the best so far is abandoned before
                // the timed launches, and a bin size whose first launch takes
                // more than four times the best is given up.
                Counters before;
                HIP_TRY(hipMemcpy(&before, b->ctx->counters, sizeof(before), hipMemcpyDeviceToHost));
                float probe = 0.0f;
                rc = timed(1, probe);
                if (rc != GLIMP_OK) break;
                Counters after;
                DrawScratch ds;
                HIP_TRY(hipMemcpy(&after, b->ctx->counters, sizeof(after), hipMemcpyDeviceToHost));
                HIP_TRY(hipMemcpy(&ds, b->scratch.scratch, sizeof(ds), hipMemcpyDeviceToHost));
                maxCount = max(maxCount, ds.max_bin_count);
                const bool slowSize = firstOfSize && found && probe > 4.0f * bestMs;
                firstOfSize = false;
                if (slowSize) {
                    skipSize = true;
                    break;
                }
                if (after.dropped != before.dropped) continue;
                if (found && probe > 2.0f * bestMs) continue;
                float ms = 0.0f;
                rc = timed(1, ms);                 // warm-up
                if (rc == GLIMP_OK) rc = timed(5, ms);
                if (rc != GLIMP_OK) break;
                if (ms < bestMs) {
                    bestMs = ms;
                    best = sh;
                    found = true;
                }
            }
            if (rc != GLIMP_OK) break;
        }
        if (rc != GLIMP_OK) break;
    }
    b->ctx->prof.enabled = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != GLIMP_OK) return rc;
    if (!found) {
        // every shape dropped something: keep the largest capacity at the
        // batch's bin size
        best.bin_capacity = caps[0];
        if (!valid_shape(best)) best.raster_threads = 256;
    }
    rc = glimp_batch_set_shape(b, &best);
    if (rc != GLIMP_OK) return rc;
    if (!found) {
        float ms = 0.0f;
        rc = timed(1, ms);
        if (rc != GLIMP_OK) return rc;
        bestMs = ms;
    }
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = bestMs;
    return GLIMP_OK;
}

}  // extern "C"

namespace {

int cull_common(glimp_context* c, glimp_scene* s, uint32_t first, uint32_t count, const float planes[20],
                const glimp_occlusion* occ, uint16_t* bitmap, uint32_t* visible, void* stream,
                const std::string& who) {
    if (!c || !s || !planes) return fail(GLIMP_E_INVALID, who + ": NULL argument");
    if ((uint64_t)first + count > s->m.count) return fail(GLIMP_E_INVALID, who + ": range outside the scene");
    Occlusion o;
    if (occ) {
        int rc = make_occlusion(occ, who, o);
        if (rc != GLIMP_OK) return rc;
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    size_t words = (count + 15) / 16;
    int rc = grow_one(c->bitmap, c->bitmap_cap, words);
    if (rc != GLIMP_OK) return rc;
    HIP_TRY(hipMemsetAsync(c->visible, 0, 4, st));
    if (count) HIP_TRY(launch_cull(s->m, first, count, planes, occ ? &o : nullptr, c->approx(), c->bitmap, c->visible, st));
    if (bitmap && words) HIP_TRY(hipMemcpyAsync(bitmap, c->bitmap, words * 2, hipMemcpyDeviceToHost, st));
    uint32_t v = 0;
    HIP_TRY(hipMemcpyAsync(&v, c->visible, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (visible) *visible = v;
    return GLIMP_OK;
}

}  // namespace

extern "C" {

int glimp_cull_meshlets(glimp_context* c, glimp_scene* s, uint32_t first, uint32_t count, const float planes[20],
                        uint16_t* bitmap, uint32_t* visible, void* stream) {
    Zone zone("glimp_cull_meshlets");
    return cull_common(c, s, first, count, planes, nullptr, bitmap, visible, stream, "glimp_cull_meshlets");
}

int glimp_cull_meshlets_occlusion(glimp_context* c, glimp_scene* s, uint32_t first, uint32_t count,
                                  const float planes[20], const glimp_occlusion* occ, uint16_t* bitmap,
                                  uint32_t* visible, void* stream) {
    if (!occ) return fail(GLIMP_E_INVALID, "glimp_cull_meshlets_occlusion: NULL occlusion");
    return cull_common(c, s, first, count, planes, occ, bitmap, visible, stream, "glimp_cull_meshlets_occlusion");
}

int glimp_hiz_create(glimp_context* c, uint32_t width, uint32_t height, glimp_hiz** out) {
    if (!c || !out) return fail(GLIMP_E_INVALID, "glimp_hiz_create: NULL argument");
    *out = nullptr;
    if (width == 0 || height == 0 || width % 4 || height % 4 || width > GLIMP_MAX_RENDER_SIZE ||
        height > GLIMP_MAX_RENDER_SIZE)
        return fail(GLIMP_E_INVALID, "glimp_hiz_create: sizes must be multiples of 4 up to GLIMP_MAX_RENDER_SIZE");
    glimp_hiz* h = new (std::nothrow) glimp_hiz();
    if (!h) return fail(GLIMP_E_NOMEM, "glimp_hiz_create: out of host memory");
    h->ctx = c;
    h->floats = hiz_layout(width, height, h->layout);
    if (h->layout.mip_levels < 2) {
        delete h;
        return fail(GLIMP_E_INVALID, "glimp_hiz_create: framebuffer too small for a depth map (needs 2 levels)");
    }
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = dev_alloc(&h->data, h->floats);
    if (e != hipSuccess) {
        delete h;
        return fail(GLIMP_E_HIP, std::string("glimp_hiz_create: ") + hipGetErrorString(e));
    }
    h->layout.depth_map = h->data;
    *out = h;
    return GLIMP_OK;
}

void glimp_hiz_destroy(glimp_hiz* h) {
    if (!h) return;
    (void)hipSetDevice(h->ctx->device);
    if (h->data) (void)hipFree(h->data);
    delete h;
}

int glimp_hiz_build(glimp_hiz* h, glimp_framebuffer* f, void* stream) {
    Zone zone("glimp_hiz_build");
    if (!h || !f) return fail(GLIMP_E_INVALID, "glimp_hiz_build: NULL argument");
    if (f->width != h->layout.width || f->height != h->layout.height)
        return fail(GLIMP_E_INVALID, "glimp_hiz_build: framebuffer and depth map sizes differ");
    if (f->content == (int)GLIMP_PIPELINE_OVERDRAW)
        return fail(GLIMP_E_UNSUPPORTED, "glimp_hiz_build: an overdraw framebuffer has no depth-test depth");
    HIP_TRY(hipSetDevice(h->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(flush_clear(f, s));
    HIP_TRY(launch_hiz_build(f->keys, h->layout, h->data, s));
    return GLIMP_OK;
}

int glimp_hiz_read(glimp_hiz* h, float* out, uint32_t* levels, uint64_t* floats, void* stream) {
    if (!h) return fail(GLIMP_E_INVALID, "glimp_hiz_read: NULL depth map");
    if (levels) *levels = h->layout.mip_levels;
    if (floats) *floats = h->floats;
    if (!out) return GLIMP_OK;
    HIP_TRY(hipSetDevice(h->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(out, h->data, h->floats * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GLIMP_OK;
}

}  // extern "C"

namespace {

int32_t cvt_trunc(float x) {   // float to int toward zero; out of range and NaN give INT32_MIN
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}
uint32_t pack_byte(float v) {  // one channel of RGBA8u::Pack: RNE, saturated to [0, 255]
    const int32_t i = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)std::nearbyint(v) : INT32_MIN;
    return (uint32_t)std::min(std::max(i, 0), 255);
}

// Sprite of a point or spot light (Shading.cpp:692-706) in the operation
// order of the reference's optimised build: clip position as one FMA chain
// from the translation column, divisions by w as a multiply by 1/w, `/ 30`
// as a multiply by 1/30; lights outside the clip volume get none.
void place_sprite(const float* m, uint32_t W, uint32_t H, ResolveLight& L) {
    L.sprite = 0;
    float c[4];
    for (int r = 0; r < 4; r++)
        c[r] = std::fma(m[8 + r], L.position[2], std::fma(m[4 + r], L.position[1], std::fma(m[r], L.position[0], m[12 + r])));
    const float mx = std::fmax(std::fmax(std::fabs(c[0]), std::fabs(c[1])), std::fabs(c[2]));
    if (mx > c[3]) return;
    const float rw = 1.0f / c[3];
    const float radius = ((float)std::max(W, H) * (1.0f / 30.0f)) * rw;
    const float cx = std::fma(c[0] * 0.5f, rw, 0.5f) * (float)W;
    const float cy = std::fma(c[1] * 0.5f, rw, 0.5f) * (float)H;
    L.depth = c[2] * rw;
    L.tile_x0 = std::max(cvt_trunc(cx - radius), 0) & ~3;
    L.tile_y0 = std::max(cvt_trunc(cy - radius), 0) & ~3;
    L.tile_x1 = std::min(cvt_trunc(cx + radius), (int32_t)W);
    L.tile_y1 = std::min(cvt_trunc(cy + radius), (int32_t)H);
    L.half_x = 0.5f - cx;
    L.half_y = 0.5f - cy;
    L.r2 = radius * radius;
    L.inv_r2 = 1.0f / L.r2;
    L.rgb = pack_byte(L.color[0] * 255.0f) | (pack_byte(L.color[1] * 255.0f) << 8) |
            (pack_byte(L.color[2] * 255.0f) << 16);
    L.sprite = 1;
}

// The launch of a resolve of framebuffer f as it stands (its clear applied):
// checks the description, allocates the colour buffer, uploads the draw
// table and fills the kernel parameters. Launches nothing.
int build_resolve(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_resolve_desc* d, hipStream_t s,
                  ResolveParams& p, const std::string& who) {
    if (d->mode > GLIMP_RESOLVE_TRIANGLE_ID) return fail(GLIMP_E_INVALID, who + ": unknown mode");
    if (d->num_lights > GLIMP_MAX_LIGHTS || (d->num_lights && !d->lights))
        return fail(GLIMP_E_LIMIT, who + ": at most GLIMP_MAX_LIGHTS lights");
    if (d->mode == GLIMP_RESOLVE_LIT && d->skybox && d->skybox->format != GLIMP_TEXEL_R11G11B10F)
        return fail(GLIMP_E_INVALID, who + ": the skybox must be an R11G11B10f texture");
    if (!f->color) HIP_TRY(dev_alloc(&f->color, (size_t)f->width * f->height));
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
    static_assert(sizeof(glimp_light) == kLightAbiBytes && offsetof(ResolveLight, sprite) == kLightAbiBytes,
                  "ResolveLight starts with glimp_light");
    for (uint32_t i = 0; i < d->num_lights; i++) {
        memcpy(&p.lights[i], &d->lights[i], sizeof(glimp_light));
        if (d->mode == GLIMP_RESOLVE_LIT && (d->flags & GLIMP_RESOLVE_LIGHT_SPRITES) && d->lights[i].type != 0)
            place_sprite(d->world_to_clip, f->width, f->height, p.lights[i]);
        p.sprites |= p.lights[i].sprite;
    }
    if (d->mode == GLIMP_RESOLVE_LIT && d->skybox) {
        p.sky = d->skybox->hdr;
        p.has_sky = 1;
    }
    p.mode = (int)d->mode;
    p.out = f->color;
    p.linear = 1;
    return GLIMP_OK;
}

}  // namespace

extern "C" {

int glimp_resolve(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_resolve_desc* d, void* stream) {
    Zone zone("glimp_resolve");
    if (!c || !f || !sc || !d) return fail(GLIMP_E_INVALID, "glimp_resolve: NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const bool od_mode = d->mode == GLIMP_RESOLVE_OVERDRAW_PIXEL || d->mode == GLIMP_RESOLVE_OVERDRAW_QUAD;
    const bool od_fb = f->content == (int)GLIMP_PIPELINE_OVERDRAW;
    if (f->content == (int)GLIMP_PIPELINE_DEFERRED)
        return fail(GLIMP_E_INVALID, "glimp_resolve: a deferred framebuffer holds G-buffer layers, not surface ids");
    if (od_mode != od_fb)
        return fail(GLIMP_E_INVALID, od_mode ? "glimp_resolve: the overdraw layers need an overdraw framebuffer"
                                             : "glimp_resolve: an overdraw framebuffer has only the overdraw layers");
    if (od_mode) {
        if (!f->color) HIP_TRY(dev_alloc(&f->color, (size_t)f->width * f->height));
        HIP_TRY(launch_resolve_overdraw(f->keys, f->aux, (int)f->width, (int)f->height, f->clear_color,
                                        d->mode == GLIMP_RESOLVE_OVERDRAW_QUAD, f->color, s, &c->prof));
        return GLIMP_OK;
    }
    HIP_TRY(flush_clear(f, s));
    ResolveParams p;
    int rc = build_resolve(c, f, sc, d, s, p, "glimp_resolve");
    if (rc != GLIMP_OK) return rc;
    HIP_TRY(launch_resolve(p, s, &c->prof));
    return GLIMP_OK;
}

}  // extern "C"

namespace {

// Shared set-up of the differentiable resolve calls: parameter arrays and
// the texture table.
int diff_common(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d, hipStream_t s, const std::string& who,
                DiffParams& p, uint64_t& num_vertices) {
    if (!c || !sc || !d) return fail(GLIMP_E_INVALID, who + ": NULL argument");
    if (!d->texels || !d->uv || !d->color || (d->num_textures && !d->textures))
        return fail(GLIMP_E_INVALID, who + ": NULL parameter array");
    if (d->num_textures > GLIMP_MAX_DIFF_TEXTURES) return fail(GLIMP_E_LIMIT, who + ": too many textures");
    if (d->num_texels >= 0xFFFFFFFFull) return fail(GLIMP_E_LIMIT, who + ": texel indices must fit 32 bits");
    num_vertices = (uint64_t)sc->m.count * 64;
    if (num_vertices >= 0xFFFFFFFFull) return fail(GLIMP_E_LIMIT, who + ": too many vertices");
    std::vector<DiffTex> tex(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; i++) {
        const glimp_diff_texture& t = d->textures[i];
        if (t.width == 0 || t.height == 0 || t.mip_levels == 0 || t.mip_levels > 16)
            return fail(GLIMP_E_INVALID, who + ": bad texture size or level count");
        for (uint32_t m = 0; m < t.mip_levels; m++) {
            const uint64_t w = t.width >> m, h = t.height >> m;
            if (w == 0 || h == 0 || t.level_offset[m] + w * h > d->num_texels)
                return fail(GLIMP_E_INVALID, who + ": texture level outside the texel array");
        }
        DiffTex& D = tex[i];
        memset(&D, 0, sizeof(D));
        D.width = t.width;
        D.height = t.height;
        D.levels = t.mip_levels;
        for (uint32_t m = 0; m < 16; m++) D.level_offset[m] = t.level_offset[m];
    }
    HIP_TRY(hipSetDevice(c->device));
    const bool same = tex.size() == c->diff_tex.size() &&
                      (tex.empty() || memcmp(tex.data(), c->diff_tex.data(), tex.size() * sizeof(DiffTex)) == 0);
    if (!same) {
        int rc = grow_one(c->d_diff_tex, c->d_diff_tex_cap, std::max<size_t>(tex.size(), 1));
        if (rc != GLIMP_OK) return rc;
        if (!tex.empty()) {
            HIP_TRY(hipStreamSynchronize(s));
            HIP_TRY(hipMemcpy(c->d_diff_tex, tex.data(), tex.size() * sizeof(DiffTex), hipMemcpyHostToDevice));
        }
        c->diff_tex = tex;
    }
    memset(&p, 0, sizeof(p));
    p.m = sc->m;
    p.materials = sc->materials;
    p.num_materials = sc->num_materials;
    p.textures = c->d_diff_tex;
    p.num_textures = d->num_textures;
    p.tables = c->approx();
    p.texels = (const float4*)d->texels;
    p.uv = (const float2*)d->uv;
    p.color = (const float4*)d->color;
    return GLIMP_OK;
}

// One framebuffer's entry of the view table.
int diff_view(glimp_framebuffer* f, const float* object_to_clip, uint32_t pixel_base, hipStream_t s,
              const std::string& who, DiffView& V) {
    if (!f) return fail(GLIMP_E_INVALID, who + ": NULL framebuffer");
    HIP_TRY(flush_clear(f, s));
    memset(&V, 0, sizeof(V));
    V.keys = f->keys;
    V.width = (int)f->width;
    V.height = (int)f->height;
    uint32_t db;
    memcpy(&db, &f->clear_depth, 4);
    V.clear_key = (unsigned long long)db << 32;
    int rc = upload_draws(f, s, &V.num_draws);
    if (rc != GLIMP_OK) return rc;
    V.draws = f->d_draws;
    V.pixel_base = pixel_base;
    memcpy(V.object_to_clip, object_to_clip, sizeof(V.object_to_clip));
    const float sw = (float)f->width, sh = (float)f->height;
    V.scale_x = 2.0f / sw;
    V.scale_y = 2.0f / sh;
    V.center_x = std::fma(0.5f, V.scale_x, -1.0f);
    V.center_y = std::fma(0.5f, V.scale_y, -1.0f);
    V.kx = 2.0f / sw;
    V.ky = -2.0f / sh;
    return GLIMP_OK;
}

// Parameters of a call over `count` views (one view is passed inline).
int diff_setup(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d, const glimp_diff_view* views,
               uint32_t count, hipStream_t s, const std::string& who, DiffParams& p, uint64_t& nv) {
    int rc = diff_common(c, sc, d, s, who, p, nv);
    if (rc != GLIMP_OK) return rc;
    if (count == 0 || !views) return fail(GLIMP_E_INVALID, who + ": no views");
    std::vector<DiffView> table(count);
    uint64_t base = 0;
    for (uint32_t i = 0; i < count; i++) {
        rc = diff_view(views[i].framebuffer, views[i].object_to_clip, (uint32_t)base, s, who, table[i]);
        if (rc != GLIMP_OK) return rc;
        base += (uint64_t)table[i].width * (uint64_t)table[i].height;
        if (base >= (1ull << 31)) return fail(GLIMP_E_LIMIT, who + ": more than 2^31 pixels in one call");
    }
    p.num_views = count;
    p.total_pixels = (uint32_t)base;
    if (count == 1) {
        p.view = table[0];
        p.views = nullptr;
    } else {
        rc = grow_one(c->d_diff_views, c->d_diff_views_cap, count);
        if (rc != GLIMP_OK) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(c->d_diff_views, table.data(), count * sizeof(DiffView), hipMemcpyHostToDevice));
        p.views = c->d_diff_views;
    }
    return GLIMP_OK;
}

int diff_forward(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d, const glimp_diff_view* views,
                 uint32_t count, float* out, void* stream, const std::string& who) {
    hipStream_t s = (hipStream_t)stream;
    DiffParams p;
    uint64_t nv = 0;
    int rc = diff_setup(c, sc, d, views, count, s, who, p, nv);
    if (rc != GLIMP_OK) return rc;
    if (!out) return fail(GLIMP_E_INVALID, who + ": NULL output");
    p.out = (float4*)out;
    HIP_TRY(launch_diff_forward(p, s, &c->prof));
    return GLIMP_OK;
}

int diff_backward(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d, const glimp_diff_view* views,
                  uint32_t count, const float* grad_out, float* grad_texels, float* grad_uv, float* grad_color,
                  void* stream, const std::string& who) {
    hipStream_t s = (hipStream_t)stream;
    DiffParams p;
    uint64_t nv = 0;
    int rc = diff_setup(c, sc, d, views, count, s, who, p, nv);
    if (rc != GLIMP_OK) return rc;
    if (!grad_out || !grad_texels || !grad_uv || !grad_color)
        return fail(GLIMP_E_INVALID, who + ": NULL gradient array");
    p.grad_out = (const float4*)grad_out;
    HIP_TRY(diff_workspace_reserve(c->diff_ws, p.total_pixels));
    HIP_TRY(launch_diff_backward(p, c->diff_ws, d->num_texels, (uint32_t)nv, (float4*)grad_texels, (float2*)grad_uv,
                                 (float4*)grad_color, s, &c->prof));
    return GLIMP_OK;
}

}  // namespace

extern "C" {

int glimp_diff_forward(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_diff_desc* d, float* out,
                       void* stream) {
    Zone zone("glimp_diff_forward");
    if (!d) return fail(GLIMP_E_INVALID, "glimp_diff_forward: NULL argument");
    glimp_diff_view v;
    v.framebuffer = f;
    memcpy(v.object_to_clip, d->object_to_clip, sizeof(v.object_to_clip));
    return diff_forward(c, sc, d, &v, 1, out, stream, "glimp_diff_forward");
}

int glimp_diff_backward(glimp_context* c, glimp_framebuffer* f, glimp_scene* sc, const glimp_diff_desc* d,
                        const float* grad_out, float* grad_texels, float* grad_uv, float* grad_color, void* stream) {
    Zone zone("glimp_diff_backward");
    if (!d) return fail(GLIMP_E_INVALID, "glimp_diff_backward: NULL argument");
    glimp_diff_view v;
    v.framebuffer = f;
    memcpy(v.object_to_clip, d->object_to_clip, sizeof(v.object_to_clip));
    return diff_backward(c, sc, d, &v, 1, grad_out, grad_texels, grad_uv, grad_color, stream, "glimp_diff_backward");
}

int glimp_diff_forward_views(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d,
                             const glimp_diff_view* views, uint32_t count, float* out, void* stream) {
    return diff_forward(c, sc, d, views, count, out, stream, "glimp_diff_forward_views");
}

int glimp_diff_backward_views(glimp_context* c, glimp_scene* sc, const glimp_diff_desc* d,
                              const glimp_diff_view* views, uint32_t count, const float* grad_out,
                              float* grad_texels, float* grad_uv, float* grad_color, void* stream) {
    return diff_backward(c, sc, d, views, count, grad_out, grad_texels, grad_uv, grad_color, stream,
                         "glimp_diff_backward_views");
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
    out->dropped = h.dropped;
    out->tile_lanes = h.tile_lanes;
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
