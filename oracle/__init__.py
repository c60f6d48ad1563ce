"""ctypes binding of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP path is compared
against. Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg may import this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")
PROBE_PATH = os.path.join(HERE, "_build", "libfastmath_probe.so")
TABLES_PATH = os.path.join(HERE, "..", "glimpsw_amd", "data", "approx14.bin")

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class Meshlets(C.Structure):
    _fields_ = [("count", C.c_uint32), ("bounds", f32p), ("num_vertices", u8p), ("num_triangles", u8p),
                ("material_id", u32p), ("tangent_handedness", u64p), ("positions", f32p),
                ("texcoords", u32p), ("normal_tangents", u32p), ("indices", u8p)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mip_levels", C.c_uint32),
                ("num_layers", C.c_uint32), ("row_shift", C.c_uint32), ("layer_stride", C.c_uint32),
                ("scale_u", C.c_float), ("scale_v", C.c_float), ("scale_lerp_u", C.c_float),
                ("scale_lerp_v", C.c_float), ("mask_u", C.c_int32), ("mask_v", C.c_int32),
                ("mask_lerp_u", C.c_int32), ("mask_lerp_v", C.c_int32),
                ("mip_offsets", C.c_uint32 * 16), ("data", u32p)]


class Material(C.Structure):
    _fields_ = [("texture", C.c_int32), ("double_sided", C.c_uint8), ("alpha_cutoff", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


class ShadingCtx(C.Structure):
    _fields_ = [("meshlets", C.POINTER(Meshlets)), ("materials", C.POINTER(Material)),
                ("textures", C.POINTER(Texture)), ("meshlet_offset", C.c_uint32),
                ("cull_bitmap", u16p), ("object_to_clip", C.c_float * 16), ("object_to_world", C.c_float * 9)]


class Framebuffer(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("layer_stride", C.c_uint32),
                ("num_layers", C.c_uint32), ("data", u32p)]


class RasterOpts(C.Structure):
    _fields_ = [("enable_clipping", C.c_uint8), ("enable_guardband", C.c_uint8), ("pipeline", C.c_uint8),
                ("pad", C.c_uint8), ("coverage_count", u32p)]


PIPELINE_VISBUFFER, PIPELINE_DEFERRED, PIPELINE_OVERDRAW = 0, 1, 2


class Counters(C.Structure):
    _fields_ = [("triangles_processed", C.c_uint64), ("triangles_rasterized", C.c_uint64),
                ("triangles_clipped", C.c_uint64), ("fragments_covered", C.c_uint64),
                ("fragments_passed", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("position", C.c_float * 3), ("direction", C.c_float * 3),
                ("color", C.c_float * 3), ("intensity", C.c_float), ("inv_radius_sq", C.c_float),
                ("spot_scale", C.c_float), ("spot_offset", C.c_float)]


class ResolveParams(C.Structure):
    _fields_ = [("inv_screen_proj", C.c_float * 16), ("object_to_clip", C.c_float * 16),
                ("object_to_world", C.c_float * 9), ("view_pos", C.c_float * 3), ("exposure", C.c_float),
                ("lights", C.POINTER(Light)), ("num_lights", C.c_uint32), ("light_sprites", C.c_uint32),
                ("world_to_clip", C.c_float * 16), ("skybox", C.POINTER(Texture))]


_lib = None


def _ptr(a: np.ndarray, t):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def _stale(target: str, *sources: str) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(os.path.join(HERE, s)) > t for s in sources)


def build(force: bool = False) -> None:
    if force or _stale(LIB_PATH, "oracle.cpp", "oracle.h"):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_rcp14.restype = C.c_float
        L.orc_rcp14.argtypes = [C.c_float]
        L.orc_rsqrt14.restype = C.c_float
        L.orc_rsqrt14.argtypes = [C.c_float]
        L.orc_div.restype = C.c_float
        L.orc_div.argtypes = [C.c_float, C.c_float]
        L.orc_cvt_rne.restype = C.c_int32
        L.orc_cvt_rne.argtypes = [C.c_float]
        L.orc_cull_meshlets.restype = C.c_uint32
        L.orc_draw_meshlets.argtypes = [C.POINTER(Framebuffer), C.c_uint32, C.POINTER(ShadingCtx),
                                        C.POINTER(RasterOpts), C.POINTER(Counters)]
        tables = np.fromfile(TABLES_PATH, dtype=np.uint16)
        assert tables.size == 131072
        L.orc_load_tables(_ptr(tables, u16p))
        _lib = L
    return _lib


def probe_lib():
    """The fast-math probe library (clang, reference flags); needs AVX-512."""
    if _stale(PROBE_PATH, os.path.join("probes", "fastmath_probe.cpp")):
        subprocess.run(["make", "-s", "-C", HERE, "probe"], check=True)
    return C.CDLL(PROBE_PATH)


class SceneBinding:
    """Keeps numpy arrays alive and exposes them as oracle structs."""

    def __init__(self, scene):
        self.scene = scene
        s = scene
        self._keep = [s.bounds, s.num_vertices, s.num_triangles, s.material_id, s.tangent_handedness,
                      s.positions, s.texcoords, s.normal_tangents, s.indices]
        self.meshlets = Meshlets(s.count, _ptr(s.bounds, f32p), _ptr(s.num_vertices, u8p),
                                 _ptr(s.num_triangles, u8p), _ptr(s.material_id, u32p),
                                 _ptr(s.tangent_handedness, u64p), _ptr(s.positions, f32p),
                                 _ptr(s.texcoords, u32p), _ptr(s.normal_tangents, u32p), _ptr(s.indices, u8p))
        self.textures = (Texture * max(1, len(s.textures)))()
        for i, t in enumerate(s.textures):
            self.textures[i] = texture_struct(t)
        self.materials = (Material * max(1, len(s.materials)))()
        for i, (tex, ds, ac) in enumerate(s.materials):
            self.materials[i] = Material(tex, int(ds), ac)

    def ctx(self, object_to_clip, meshlet_offset=0, cull_bitmap=None, object_to_world=None):
        c = ShadingCtx()
        o2w = np.eye(3).reshape(9) if object_to_world is None else np.asarray(object_to_world).reshape(9)
        for i in range(9):
            c.object_to_world[i] = float(o2w[i])
        c.meshlets = C.pointer(self.meshlets)
        c.materials = C.cast(self.materials, C.POINTER(Material))
        c.textures = C.cast(self.textures, C.POINTER(Texture))
        c.meshlet_offset = meshlet_offset
        c.cull_bitmap = _ptr(cull_bitmap, u16p) if cull_bitmap is not None else u16p()
        for i in range(16):
            c.object_to_clip[i] = float(object_to_clip[i])
        return c


def texture_struct(t) -> Texture:
    T = Texture()
    T.width, T.height, T.mip_levels, T.num_layers = t.width, t.height, t.mip_levels, t.num_layers
    T.row_shift, T.layer_stride = t.row_shift, t.layer_stride
    T.scale_u, T.scale_v = float(t.width), float(t.height)
    T.mask_u, T.mask_v = t.width - 1, t.height - 1
    T.mask_lerp_u, T.mask_lerp_v = t.mask_lerp_u, t.mask_lerp_v
    T.scale_lerp_u, T.scale_lerp_v = float(t.mask_lerp_u + 1), float(t.mask_lerp_v + 1)
    for i in range(16):
        T.mip_offsets[i] = int(t.mip_offsets[i])
    T.data = _ptr(t.data, u32p)
    return T


class OracleFramebuffer:
    """Two-layer (surface id, depth) framebuffer in 4x4 tiles."""

    def __init__(self, width: int, height: int, layers: int = 2):
        assert width % 4 == 0 and height % 4 == 0
        self.width, self.height = width, height
        self.layer_stride = (width * height + 63) & ~63
        self.data = np.zeros(self.layer_stride * layers, dtype=np.uint32)
        self.fb = Framebuffer(width, height, self.layer_stride, layers, _ptr(self.data, u32p))

    def clear(self, color: int = 0, depth: float = 0.0):
        lib().orc_fb_clear(C.byref(self.fb), C.c_uint32(color), C.c_float(depth))

    def pixels(self, layer: int) -> np.ndarray:
        out = np.zeros((self.height, self.width), dtype=np.uint32)
        lib().orc_get_pixels(C.byref(self.fb), C.c_uint32(layer), _ptr(out, u32p), C.c_uint32(self.width))
        return out

    def ids(self) -> np.ndarray:
        return self.pixels(0)

    def depth(self) -> np.ndarray:
        return self.pixels(1).view(np.float32)


def draw(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, count: int, meshlet_offset: int = 0,
         cull_bitmap=None, clipping: bool = True, guardband: bool = True, coverage: bool = False,
         pipeline: int = PIPELINE_VISBUFFER, object_to_world=None, rows=None):
    """DrawMeshlets into `fb`; with `rows` = (row0, row1) (multiples of 4)
    only fragments in those rows are drawn and counted."""
    cov = np.zeros((fb.height, fb.width), dtype=np.uint32) if coverage else None
    opts = RasterOpts(int(clipping), int(guardband), pipeline, 0, _ptr(cov, u32p) if cov is not None else u32p())
    ctr = Counters()
    if pipeline == PIPELINE_DEFERRED:
        assert fb.fb.num_layers >= 3, "the deferred pipeline writes three layers"
    ctx = binding.ctx(object_to_clip, meshlet_offset, cull_bitmap, object_to_world)
    r0, r1 = rows if rows is not None else (0, fb.height)
    lib().orc_draw_meshlets_rows(C.byref(fb.fb), C.c_uint32(count), C.byref(ctx), C.byref(opts), C.byref(ctr),
                                 C.c_uint32(r0), C.c_uint32(r1))
    return ctr.as_dict(), cov


def cull(binding: SceneBinding, planes: np.ndarray, first: int, count: int):
    bitmap = np.zeros((count + 15) // 16, dtype=np.uint16)
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    n = lib().orc_cull_meshlets(_ptr(bitmap, u16p), C.byref(binding.meshlets), C.c_uint32(first),
                                C.c_uint32(count), _ptr(planes, f32p))
    return bitmap, int(n)


class Occlusion(C.Structure):
    _fields_ = [("depth_map", f32p), ("mip_levels", C.c_uint32), ("object_to_prev_view", C.c_float * 16),
                ("object_to_world_scale", C.c_float), ("proj", C.c_float * 3), ("frame_size", C.c_float * 2)]


def hiz_levels(width: int, height: int) -> int:
    """Mip levels of the reference's depth map for a framebuffer."""
    lib().orc_hiz_levels.restype = C.c_uint32
    return int(lib().orc_hiz_levels(C.c_uint32(width), C.c_uint32(height)))


def hiz_level_dims(width: int, height: int):
    return [(((width - 1) >> (m + 1)) + 1, ((height - 1) >> (m + 1)) + 1)
            for m in range(max(hiz_levels(width, height) - 1, 0))]


def hiz_build(fb: OracleFramebuffer) -> np.ndarray:
    """DownsampleDepth: the min-depth pyramid, levels concatenated row-major."""
    n = sum(w * h for w, h in hiz_level_dims(fb.width, fb.height))
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_hiz_build(C.byref(fb.fb), _ptr(out, f32p))
    return out[:n]


def occlusion(depth_map: np.ndarray, width: int, height: int, object_to_prev_view, scale: float, proj, keep: list):
    o = Occlusion()
    dm = np.ascontiguousarray(depth_map, np.float32)
    keep.append(dm)
    o.depth_map = _ptr(dm, f32p)
    o.mip_levels = hiz_levels(width, height)
    for i in range(16):
        o.object_to_prev_view[i] = float(object_to_prev_view[i])
    o.object_to_world_scale = float(scale)
    for i in range(3):
        o.proj[i] = float(proj[i])
    o.frame_size[0], o.frame_size[1] = float(width), float(height)
    return o


def cull_occlusion(binding: SceneBinding, planes: np.ndarray, first: int, count: int, occ: Occlusion):
    bitmap = np.zeros((count + 15) // 16, dtype=np.uint16)
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    lib().orc_cull_meshlets_occlusion.restype = C.c_uint32
    n = lib().orc_cull_meshlets_occlusion(_ptr(bitmap, u16p), C.byref(binding.meshlets), C.c_uint32(first),
                                          C.c_uint32(count), _ptr(planes, f32p), C.byref(occ))
    return bitmap, int(n)


def resolve_albedo(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, meshlet_offset: int = 0):
    ctx = binding.ctx(object_to_clip, meshlet_offset)
    lib().orc_resolve_albedo(C.byref(fb.fb), C.byref(ctx))
    return fb.pixels(0)


(DEBUG_BASE_COLOR, DEBUG_NORMALS, DEBUG_METALLIC_ROUGHNESS, DEBUG_MESHLET_ID, DEBUG_TRIANGLE_ID,
 DEBUG_OVERDRAW_PIXEL, DEBUG_OVERDRAW_QUAD) = range(7)


def resolve_debug(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, layer: int,
                  object_to_world=None, meshlet_offset: int = 0):
    """ResolveDebug of one layer into layer 0; returns row-major RGBA8."""
    ctx = binding.ctx(object_to_clip, meshlet_offset)
    o2w = np.ascontiguousarray(object_to_world if object_to_world is not None else np.eye(3).reshape(9), np.float32)
    lib().orc_resolve_debug(C.byref(fb.fb), C.byref(ctx), C.c_uint32(layer), _ptr(o2w, f32p))
    return fb.pixels(0)


def render_frame(scene, view, binding=None, cull_first: bool = True, clipping: bool = True):
    """Clear, cull, draw: the reference frame sequence (Main.cpp:213-240)."""
    binding = binding or SceneBinding(scene)
    fb = OracleFramebuffer(view.width, view.height)
    fb.clear(0, 0.0)
    bitmap = None
    if cull_first:
        bitmap, _ = cull(binding, view.frustum_planes(), view.meshlet_offset, view.meshlet_count)
    ctr, _ = draw(fb, binding, view.object_to_clip, view.meshlet_count, view.meshlet_offset, bitmap,
                  clipping=clipping)
    return fb, ctr, binding


def resolve_lit(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, lighting, meshlet_offset: int = 0,
                rows=None):
    """Resolve (lit, tonemapped) into layer 0; returns row-major RGBA8. With
    `rows` = (row0, row1) only those rows are resolved and nothing is read
    back (returns None)."""
    from glimpsw_amd import light_structs

    ctx = binding.ctx(object_to_clip, meshlet_offset)
    rp = ResolveParams()
    for i in range(16):
        rp.inv_screen_proj[i] = float(lighting.inv_screen_proj[i])
        rp.object_to_clip[i] = float(object_to_clip[i])
    for i in range(9):
        rp.object_to_world[i] = float(lighting.object_to_world[i])
    for i in range(3):
        rp.view_pos[i] = float(lighting.view_pos[i])
    rp.exposure = lighting.exposure
    lights = light_structs(lighting.lights, Light)
    rp.lights = C.cast(lights, C.POINTER(Light))
    rp.num_lights = len(lighting.lights)
    if getattr(lighting, "light_sprites", False):
        rp.light_sprites = 1
        for i in range(16):
            rp.world_to_clip[i] = float(lighting.world_to_clip[i])
    sky = getattr(lighting, "skybox", None)
    if sky is not None:
        sky_struct = texture_struct(sky)
        rp.skybox = C.pointer(sky_struct)
    if rows is not None:
        lib().orc_resolve_lit_rows(C.byref(fb.fb), C.byref(ctx), C.byref(rp), C.c_uint32(rows[0]),
                                   C.c_uint32(rows[1]))
        return None
    lib().orc_resolve_lit(C.byref(fb.fb), C.byref(ctx), C.byref(rp))
    return fb.pixels(0)


# ---------------------------------------------------------------- differentiable resolve

f64p = C.POINTER(C.c_double)


class DiffTexture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mip_levels", C.c_uint32), ("pad", C.c_uint32),
                ("level_offset", C.c_uint64 * 16)]


class DiffParams(C.Structure):
    _fields_ = [("object_to_clip", C.c_float * 16), ("textures", C.POINTER(DiffTexture)),
                ("num_textures", C.c_uint32), ("texels", f32p), ("uv", f32p), ("color", f32p),
                ("texels_d", f64p), ("uv_d", f64p), ("color_d", f64p)]


def _diff_params(chains, object_to_clip, texels, uv, color, texels_d=None, uv_d=None, color_d=None):
    keep = []
    tex = (DiffTexture * max(1, len(chains.sizes)))()
    for i, ((w, h, L), offs) in enumerate(zip(chains.sizes, chains.offsets)):
        tex[i].width, tex[i].height, tex[i].mip_levels = w, h, L
        for m, o in enumerate(offs):
            tex[i].level_offset[m] = o
    p = DiffParams()
    for i in range(16):
        p.object_to_clip[i] = float(object_to_clip[i])
    p.textures = C.cast(tex, C.POINTER(DiffTexture))
    p.num_textures = len(chains.sizes)

    def f32(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return _ptr(a, f32p)

    def f64(a):
        if a is None:
            return f64p()
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _ptr(a, f64p)

    p.texels, p.uv, p.color = f32(texels), f32(uv), f32(color)
    p.texels_d, p.uv_d, p.color_d = f64(texels_d), f64(uv_d), f64(color_d)
    return p, (tex, keep)


def diff_forward(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, chains, texels, uv, color,
                 texels_d=None, uv_d=None, color_d=None) -> np.ndarray:
    """FP64 forward of the differentiable resolve: (H, W, 4) float64.
    Discrete choices come from the fp32 ``texels/uv/color``; the optional
    ``*_d`` float64 arrays replace their values in the continuous part."""
    ctx = binding.ctx(object_to_clip)
    p, keep = _diff_params(chains, object_to_clip, texels, uv, color, texels_d, uv_d, color_d)
    out = np.zeros((fb.height, fb.width, 4), dtype=np.float64)
    lib().orc_diff_forward(C.byref(fb.fb), C.byref(ctx), C.byref(p), _ptr(out, f64p))
    return out


def diff_backward(fb: OracleFramebuffer, binding: SceneBinding, object_to_clip, chains, texels, uv, color,
                  grad_out) -> tuple:
    """FP64 gradients of sum(grad_out * out): (texels, uv, color)."""
    ctx = binding.ctx(object_to_clip)
    p, keep = _diff_params(chains, object_to_clip, texels, uv, color)
    g = np.ascontiguousarray(grad_out, dtype=np.float64)
    gt = np.zeros((np.asarray(texels).shape[0], 4), dtype=np.float64)
    gu = np.zeros((np.asarray(uv).shape[0], 2), dtype=np.float64)
    gc = np.zeros((np.asarray(color).shape[0], 4), dtype=np.float64)
    lib().orc_diff_backward(C.byref(fb.fb), C.byref(ctx), C.byref(p), _ptr(g, f64p), _ptr(gt, f64p),
                            _ptr(gu, f64p), _ptr(gc, f64p))
    return gt, gu, gc


def bary_f64(clip, su: float, sv: float) -> np.ndarray:
    c = np.ascontiguousarray(clip, dtype=np.float32).reshape(12)
    out = np.zeros(3, dtype=np.float64)
    lib().orc_bary_f64(_ptr(c, f32p), C.c_double(su), C.c_double(sv), _ptr(out, f64p))
    return out


def row_bands(height: int, parts: int) -> list:
    """`parts` row ranges of about equal height covering [0, height), cut on
    4-row tile boundaries (empty ranges dropped)."""
    tiles = height // 4
    cuts = [4 * ((tiles * k) // parts) for k in range(parts + 1)]
    return [(a, b) for a, b in zip(cuts, cuts[1:]) if b > a]


def render_frame_threads(scene, view, binding, lighting, pool, bands):
    """The frame of render_frame + resolve_lit split over host threads by row
    bands: the cull once, then every band's draw and every band's lit resolve
    on `pool` (a concurrent.futures executor). Bit-identical to the one-thread
    frame; returns the framebuffer."""
    fb = OracleFramebuffer(view.width, view.height)
    fb.clear(0, 0.0)
    bitmap, _ = cull(binding, view.frustum_planes(), view.meshlet_offset, view.meshlet_count)
    list(pool.map(lambda r: draw(fb, binding, view.object_to_clip, view.meshlet_count, view.meshlet_offset, bitmap,
                                 rows=r), bands))
    list(pool.map(lambda r: resolve_lit(fb, binding, view.object_to_clip, lighting, view.meshlet_offset, rows=r),
                  bands))
    return fb
