"""glimpsw_amd: MI355X-native meshlet raster -> vis-buffer -> resolve path.

Python binding of the C ABI in include/glimp.h (libglimp.so, HIP, gfx950).
There is no CPU fallback: every call goes to the HIP library, and loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libglimp.so")

GLIMP_DRAW_NO_CLIPPING = 0x1
GLIMP_DRAW_NO_GUARDBAND = 0x2
GLIMP_DRAW_DIRECT = 0x4
GLIMP_DRAW_COUNTERS = 0x8
GLIMP_PIPELINE_VISBUFFER = 0
GLIMP_PIPELINE_DEFERRED = 1
GLIMP_PIPELINE_OVERDRAW = 2
MAX_RENDER_SIZE = 2896

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)


class GlimpError(RuntimeError):
    pass


class MeshletsDesc(C.Structure):
    _fields_ = [("count", C.c_uint32), ("bounds", f32p), ("num_vertices", u8p), ("num_triangles", u8p),
                ("material_id", u32p), ("tangent_handedness", u64p), ("positions", f32p),
                ("texcoords", u32p), ("normal_tangents", u32p), ("indices", u8p)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mip_levels", C.c_uint32),
                ("num_layers", C.c_uint32), ("row_shift", C.c_uint32), ("layer_stride", C.c_uint32),
                ("mip_offsets", C.c_uint32 * 16), ("texels", u32p), ("format", C.c_uint32)]


class MaterialDesc(C.Structure):
    _fields_ = [("texture", C.c_int32), ("double_sided", C.c_uint8), ("alpha_cutoff", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


class OcclusionDesc(C.Structure):
    _fields_ = [("depth_map", C.c_void_p), ("object_to_prev_view", C.c_float * 16),
                ("object_to_world_scale", C.c_float), ("proj_factors", C.c_float * 3)]


class DrawDesc(C.Structure):
    _fields_ = [("meshlet_offset", C.c_uint32), ("meshlet_count", C.c_uint32),
                ("object_to_clip", C.c_float * 16), ("frustum_planes", f32p), ("flags", C.c_uint32),
                ("pipeline", C.c_uint32), ("bin_size", C.c_uint32), ("bin_capacity", C.c_uint32),
                ("occlusion", C.POINTER(OcclusionDesc)), ("object_to_world", C.c_float * 9)]


class ViewDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("framebuffer", C.c_void_p), ("meshlet_offset", C.c_uint32),
                ("meshlet_count", C.c_uint32), ("object_to_clip", C.c_float * 16), ("frustum_planes", f32p),
                ("flags", C.c_uint32), ("clear_color", C.c_uint32), ("clear_depth", C.c_float),
                ("occlusion", C.POINTER(OcclusionDesc))]


class BatchShape(C.Structure):
    _fields_ = [("bin_size", C.c_uint32), ("bin_capacity", C.c_uint32), ("raster_threads", C.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class MeshletRange(C.Structure):
    _fields_ = [("vertex_offset", C.c_uint32), ("triangle_offset", C.c_uint32), ("vertex_count", C.c_uint32),
                ("triangle_count", C.c_uint32)]


class DiffTextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mip_levels", C.c_uint32), ("pad", C.c_uint32),
                ("level_offset", C.c_uint64 * 16)]


class DiffDesc(C.Structure):
    _fields_ = [("object_to_clip", C.c_float * 16), ("textures", C.POINTER(DiffTextureDesc)),
                ("num_textures", C.c_uint32), ("num_texels", C.c_uint64), ("texels", C.c_void_p),
                ("uv", C.c_void_p), ("color", C.c_void_p)]


class DiffViewDesc(C.Structure):
    _fields_ = [("framebuffer", C.c_void_p), ("object_to_clip", C.c_float * 16)]


class CountersStruct(C.Structure):
    _fields_ = [("triangles_processed", C.c_uint64), ("triangles_rasterized", C.c_uint64),
                ("triangles_clipped", C.c_uint64), ("bins_rasterized", C.c_uint64),
                ("fragments_covered", C.c_uint64), ("fragments_passed", C.c_uint64),
                ("setup_records", C.c_uint64), ("bin_entries", C.c_uint64), ("dropped", C.c_uint64),
                ("tile_lanes", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("position", C.c_float * 3), ("direction", C.c_float * 3),
                ("color", C.c_float * 3), ("intensity", C.c_float), ("inv_radius_sq", C.c_float),
                ("spot_scale", C.c_float), ("spot_offset", C.c_float)]


class ResolveDesc(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("object_to_clip", C.c_float * 16), ("object_to_world", C.c_float * 9),
                ("inv_screen_proj", C.c_float * 16), ("view_pos", C.c_float * 3), ("exposure", C.c_float),
                ("lights", C.POINTER(LightDesc)), ("num_lights", C.c_uint32), ("flags", C.c_uint32),
                ("world_to_clip", C.c_float * 16), ("skybox", C.c_void_p)]


GLIMP_RESOLVE_ALBEDO = 0
GLIMP_RESOLVE_LIT = 1
GLIMP_RESOLVE_NORMALS = 2
GLIMP_RESOLVE_METALLIC_ROUGHNESS = 3
GLIMP_RESOLVE_MESHLET_ID = 4
GLIMP_RESOLVE_TRIANGLE_ID = 5
GLIMP_RESOLVE_OVERDRAW_PIXEL = 6
GLIMP_RESOLVE_OVERDRAW_QUAD = 7
GLIMP_RESOLVE_LIGHT_SPRITES = 0x1


def light_structs(lights, cls=None):
    cls = cls or LightDesc
    arr = (cls * max(1, len(lights)))()
    for i, L in enumerate(lights):
        arr[i].type = L.type
        for k in range(3):
            arr[i].position[k] = L.position[k]
            arr[i].direction[k] = L.direction[k]
            arr[i].color[k] = L.color[k]
        arr[i].intensity = L.intensity
        arr[i].inv_radius_sq = L.inv_radius_sq
        arr[i].spot_scale = 0.0
        arr[i].spot_offset = 0.0
    return arr


# Every symbol include/glimp.h declares, with its ctypes signature.
_SIGNATURES = {
    "glimp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "glimp_destroy": (None, [C.c_void_p]),
    "glimp_last_error": (C.c_char_p, []),
    "glimp_scene_create": (C.c_int, [C.c_void_p, C.POINTER(MeshletsDesc), C.POINTER(C.c_void_p)]),
    "glimp_scene_destroy": (None, [C.c_void_p]),
    "glimp_texture_create": (C.c_int, [C.c_void_p, C.POINTER(TextureDesc), C.POINTER(C.c_void_p)]),
    "glimp_texture_destroy": (None, [C.c_void_p]),
    "glimp_scene_set_materials": (C.c_int, [C.c_void_p, C.POINTER(MaterialDesc), C.c_uint32,
                                            C.POINTER(C.c_void_p), C.c_uint32]),
    "glimp_framebuffer_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "glimp_framebuffer_destroy": (None, [C.c_void_p]),
    "glimp_framebuffer_clear": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]),
    "glimp_framebuffer_read": (C.c_int, [C.c_void_p, u32p, f32p, C.c_int, C.c_void_p]),
"glimp_framebuffer_read_layer": (C.c_int, [C.c_void_p, C.c_uint32, u32p, C.c_int, C.c_void_p]),
    "glimp_framebuffer_keys": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "glimp_draw_meshlets": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(DrawDesc), C.c_void_p]),
    "glimp_draw_meshlets_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(DrawDesc), C.c_void_p]),
    "glimp_cull_meshlets": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, f32p, u16p,
                                      C.POINTER(C.c_uint32), C.c_void_p]),
    "glimp_cull_meshlets_occlusion": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, f32p,
                                                C.POINTER(OcclusionDesc), u16p, C.POINTER(C.c_uint32), C.c_void_p]),
    "glimp_hiz_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "glimp_hiz_destroy": (None, [C.c_void_p]),
    "glimp_hiz_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "glimp_hiz_read": (C.c_int, [C.c_void_p, f32p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p]),
    "glimp_get_counters": (C.c_int, [C.c_void_p, C.POINTER(CountersStruct)]),
    "glimp_reset_counters": (C.c_int, [C.c_void_p]),
    "glimp_resolve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ResolveDesc), C.c_void_p]),
    "glimp_framebuffer_color": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "glimp_framebuffer_read_color": (C.c_int, [C.c_void_p, u32p, C.c_void_p]),
    "glimp_batch_create": (C.c_int, [C.c_void_p, C.POINTER(ViewDesc), C.c_uint32, C.POINTER(BatchShape),
                                     C.POINTER(C.c_void_p)]),
    "glimp_batch_destroy": (None, [C.c_void_p]),
    "glimp_batch_set_shape": (C.c_int, [C.c_void_p, C.POINTER(BatchShape)]),
    "glimp_batch_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "glimp_batch_launch_per_view": (C.c_int, [C.c_void_p, C.c_void_p]),
    "glimp_batch_set_resolve": (C.c_int, [C.c_void_p, C.POINTER(ResolveDesc)]),
    "glimp_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BatchShape), C.POINTER(C.c_float)]),
    "glimp_diff_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(DiffDesc), C.c_void_p,
                                     C.c_void_p]),
    "glimp_diff_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(DiffDesc), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "glimp_diff_forward_views": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DiffDesc), C.POINTER(DiffViewDesc),
                                           C.c_uint32, C.c_void_p, C.c_void_p]),
    "glimp_diff_backward_views": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DiffDesc), C.POINTER(DiffViewDesc),
                                            C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "glimp_build_meshlets_bound": (C.c_size_t, [C.c_size_t, C.c_size_t, C.c_size_t]),
    "glimp_build_meshlets": (C.c_int, [u32p, C.c_size_t, f32p, C.c_size_t, C.c_size_t, C.c_uint32, C.c_uint32,
                                       C.POINTER(MeshletRange), u32p, C.POINTER(C.c_uint8), C.POINTER(C.c_size_t)]),
    "glimp_meshlet_bounds": (C.c_int, [u32p, C.c_uint32, f32p, C.c_size_t, f32p]),
    "glimp_pack_normal_tangents": (C.c_int, [f32p, f32p, C.c_size_t, u32p]),
    "glimp_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "glimp_profile_read": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_uint32, C.POINTER(C.c_uint32), C.c_int]),
}

_lib = None


def library_path() -> str:
    return LIB_PATH


def load(build_if_missing: bool = False):
    """Loads libglimp.so. Raises if it is missing (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        if build_if_missing:
            from . import build as _b
            _b.build()
        else:
            raise GlimpError(f"libglimp.so not built ({LIB_PATH}); run python -m glimpsw_amd.build")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7, and a
    # second copy loaded beside it cannot share the GPU (whichever initialises
    # second finds no devices). Importing torch first makes libglimp bind to
    # the copy torch loaded (same soname), so device memory and streams are
    # shared with torch tensors.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = load().glimp_last_error().decode(errors="replace")
        raise GlimpError(f"{what} failed ({rc}): {msg}")


def _ptr(a: np.ndarray, t):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(t)


def _stream_ptr(stream) -> C.c_void_p:
    if stream is None:
        return C.c_void_p(0)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(int(stream.cuda_stream))


class Context:
    """A glimp context on one HIP device (include/glimp.h glimp_create)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.glimp_create(C.byref(h), device), "glimp_create")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self.lib.glimp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counters(self) -> dict:
        c = CountersStruct()
        _check(self.lib.glimp_get_counters(self.handle, C.byref(c)), "glimp_get_counters")
        return c.as_dict()

    def reset_counters(self):
        _check(self.lib.glimp_reset_counters(self.handle), "glimp_reset_counters")

    def profile(self, enable: bool):
        _check(self.lib.glimp_profile_enable(self.handle, int(enable)), "glimp_profile_enable")

    def profile_read(self, reset: bool = True) -> dict:
        """{kernel name: (launches, total ms)} since the last reset."""
        buf = (KernelTime * 64)()
        n = C.c_uint32()
        _check(self.lib.glimp_profile_read(self.handle, buf, 64, C.byref(n), int(reset)), "glimp_profile_read")
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(min(n.value, 64))}


class Texture:
    def __init__(self, ctx: Context, tex):
        self.ctx = ctx
        d = TextureDesc()
        d.width, d.height, d.mip_levels, d.num_layers = tex.width, tex.height, tex.mip_levels, tex.num_layers
        d.row_shift, d.layer_stride = tex.row_shift, tex.layer_stride
        for i in range(16):
            d.mip_offsets[i] = int(tex.mip_offsets[i])
        data = np.ascontiguousarray(tex.data, dtype=np.uint32)
        d.texels = _ptr(data, u32p)
        d.format = int(getattr(tex, "format", 0))
        h = C.c_void_p()
        _check(ctx.lib.glimp_texture_create(ctx.handle, C.byref(d), C.byref(h)), "glimp_texture_create")
        self.handle = h

    def close(self):
        if self.handle:
            self.ctx.lib.glimp_texture_destroy(self.handle)
            self.handle = None


class Scene:
    """Device copy of a glimpsw_amd.scene.Scene (meshlets, materials, textures)."""

    def __init__(self, ctx: Context, scene):
        self.ctx = ctx
        self.host = scene
        s = scene
        d = MeshletsDesc(s.count, _ptr(s.bounds, f32p), _ptr(s.num_vertices, u8p), _ptr(s.num_triangles, u8p),
                         _ptr(s.material_id, u32p), _ptr(s.tangent_handedness, u64p), _ptr(s.positions, f32p),
                         _ptr(s.texcoords, u32p), _ptr(s.normal_tangents, u32p), _ptr(s.indices, u8p))
        h = C.c_void_p()
        _check(ctx.lib.glimp_scene_create(ctx.handle, C.byref(d), C.byref(h)), "glimp_scene_create")
        self.handle = h
        self.textures = [Texture(ctx, t) for t in s.textures]
        mats = (MaterialDesc * max(1, len(s.materials)))()
        for i, (tex, ds, ac) in enumerate(s.materials):
            mats[i] = MaterialDesc(tex, int(ds), ac)
        texh = (C.c_void_p * max(1, len(self.textures)))(*[t.handle.value for t in self.textures])
        _check(ctx.lib.glimp_scene_set_materials(h, mats, len(s.materials), texh, len(self.textures)),
               "glimp_scene_set_materials")

    def close(self):
        if self.handle:
            self.ctx.lib.glimp_scene_destroy(self.handle)
            self.handle = None
        for t in self.textures:
            t.close()
        self.textures = []

    def cull(self, planes: np.ndarray, first: int, count: int, stream=None):
        bitmap = np.zeros(max(1, (count + 15) // 16), dtype=np.uint16)
        vis = C.c_uint32()
        planes = np.ascontiguousarray(planes, dtype=np.float32)
        _check(self.ctx.lib.glimp_cull_meshlets(self.ctx.handle, self.handle, first, count, _ptr(planes, f32p),
                                                _ptr(bitmap, u16p), C.byref(vis), _stream_ptr(stream)),
               "glimp_cull_meshlets")
        return bitmap[: (count + 15) // 16], int(vis.value)

    def cull_occlusion(self, planes: np.ndarray, first: int, count: int, occlusion: "Occlusion", stream=None):
        """CullMeshlets with the occlusion branch (HiZ test of the frustum survivors)."""
        bitmap = np.zeros(max(1, (count + 15) // 16), dtype=np.uint16)
        vis = C.c_uint32()
        planes = np.ascontiguousarray(planes, dtype=np.float32)
        _check(self.ctx.lib.glimp_cull_meshlets_occlusion(self.ctx.handle, self.handle, first, count,
                                                          _ptr(planes, f32p), C.byref(occlusion.desc),
                                                          _ptr(bitmap, u16p), C.byref(vis), _stream_ptr(stream)),
               "glimp_cull_meshlets_occlusion")
        return bitmap[: (count + 15) // 16], int(vis.value)


class DepthMap:
    """The previous frame's min-depth pyramid for occlusion culling
    (glimp_hiz_* in include/glimp.h)."""

    def __init__(self, ctx: "Context", width: int, height: int):
        self.ctx = ctx
        self.width, self.height = width, height
        h = C.c_void_p()
        _check(ctx.lib.glimp_hiz_create(ctx.handle, width, height, C.byref(h)), "glimp_hiz_create")
        self.handle = h.value

    def build(self, fb: "Framebuffer", stream=None):
        _check(self.ctx.lib.glimp_hiz_build(self.handle, fb.handle, _stream_ptr(stream)), "glimp_hiz_build")

    def read(self, stream=None):
        levels, floats = C.c_uint32(), C.c_uint64()
        _check(self.ctx.lib.glimp_hiz_read(self.handle, None, C.byref(levels), C.byref(floats), None), "glimp_hiz_read")
        out = np.zeros(max(int(floats.value), 1), np.float32)
        _check(self.ctx.lib.glimp_hiz_read(self.handle, _ptr(out, f32p), None, None, _stream_ptr(stream)),
               "glimp_hiz_read")
        return out[: int(floats.value)], int(levels.value)

    def close(self):
        if self.handle:
            self.ctx.lib.glimp_hiz_destroy(self.handle)
            self.handle = None


class Occlusion:
    """Occlusion inputs of one cull or draw: the depth map and the terms
    CullMeshlets derives from its matrices (Shading.cpp:779-795)."""

    def __init__(self, depth_map: DepthMap, object_to_prev_view, object_to_world_scale: float, proj_factors):
        self.depth_map = depth_map
        d = OcclusionDesc()
        d.depth_map = depth_map.handle
        for i in range(16):
            d.object_to_prev_view[i] = float(object_to_prev_view[i])
        d.object_to_world_scale = float(object_to_world_scale)
        for i in range(3):
            d.proj_factors[i] = float(proj_factors[i])
        self.desc = d


def resolve_desc(object_to_clip, mode: int = GLIMP_RESOLVE_ALBEDO, lighting=None, skybox: "Texture | None" = None):
    """A glimp_resolve_desc (and the light array it points to, to keep
    alive): ``lighting`` supplies the lit terms and object_to_world; its
    light sprites need ``world_to_clip``; ``skybox`` is a device
    R11G11B10f texture for the uncovered pixels of the lit resolve."""
    d = ResolveDesc()
    d.mode = mode
    if skybox is not None:
        d.skybox = skybox.handle
    for i in range(16):
        d.object_to_clip[i] = float(object_to_clip[i])
    keep = None
    if lighting is not None:
        for i in range(9):
            d.object_to_world[i] = float(lighting.object_to_world[i])
        for i in range(16):
            d.inv_screen_proj[i] = float(lighting.inv_screen_proj[i])
        for i in range(3):
            d.view_pos[i] = float(lighting.view_pos[i])
        d.exposure = lighting.exposure
        keep = light_structs(lighting.lights)
        d.lights = C.cast(keep, C.POINTER(LightDesc))
        d.num_lights = len(lighting.lights)
        if getattr(lighting, "light_sprites", False) and getattr(lighting, "world_to_clip", None) is not None:
            d.flags |= GLIMP_RESOLVE_LIGHT_SPRITES
            for i in range(16):
                d.world_to_clip[i] = float(lighting.world_to_clip[i])
    else:
        d.exposure = 1.0
        d.object_to_world[0] = d.object_to_world[4] = d.object_to_world[8] = 1.0
    return d, keep


class Framebuffer:
    """Device vis-buffer of 64-bit depth-test keys (include/glimp.h)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.lib.glimp_framebuffer_create(ctx.handle, width, height, C.byref(h)), "glimp_framebuffer_create")
        self.handle = h
        self.width, self.height = width, height

    def close(self):
        if self.handle:
            self.ctx.lib.glimp_framebuffer_destroy(self.handle)
            self.handle = None

    def clear(self, color: int = 0, depth: float = 0.0, stream=None):
        _check(self.ctx.lib.glimp_framebuffer_clear(self.handle, color, depth, _stream_ptr(stream)),
               "glimp_framebuffer_clear")

    def draw(self, scene: Scene, object_to_clip, count: int, offset: int = 0, planes=None, flags: int = 0,
             stream=None, bin_size: int = 0, bin_capacity: int = 0, occlusion: "Occlusion | None" = None,
             pipeline: int = GLIMP_PIPELINE_VISBUFFER, object_to_world=None):
        """Rasterizer::DrawMeshlets with the pipeline's dispatch table;
        ``object_to_world`` (mat3, column-major, default identity) is the
        deferred pipeline's ObjectToWorldMat."""
        d = DrawDesc()
        o2w = np.eye(3).reshape(9) if object_to_world is None else np.asarray(object_to_world).reshape(9)
        for i in range(9):
            d.object_to_world[i] = float(o2w[i])
        d.meshlet_offset, d.meshlet_count = offset, count
        for i in range(16):
            d.object_to_clip[i] = float(object_to_clip[i])
        keep = None
        if planes is not None:
            keep = np.ascontiguousarray(planes, dtype=np.float32)
            d.frustum_planes = _ptr(keep, f32p)
        d.flags = flags
        d.pipeline = pipeline
        d.bin_size, d.bin_capacity = bin_size, bin_capacity
        if occlusion is not None:
            d.occlusion = C.pointer(occlusion.desc)
        _check(self.ctx.lib.glimp_draw_meshlets(self.ctx.handle, self.handle, scene.handle, C.byref(d),
                                                _stream_ptr(stream)), "glimp_draw_meshlets")

    def read(self, linear: bool = True, stream=None):
        ids = np.zeros((self.height, self.width), dtype=np.uint32)
        depth = np.zeros((self.height, self.width), dtype=np.float32)
        _check(self.ctx.lib.glimp_framebuffer_read(self.handle, _ptr(ids, u32p), _ptr(depth, f32p), int(linear),
                                                   _stream_ptr(stream)), "glimp_framebuffer_read")
        return ids, depth

    def read_layer(self, layer: int, linear: bool = True, stream=None) -> np.ndarray:
        """One framebuffer layer as 32-bit words (glimp_framebuffer_read_layer)."""
        out = np.zeros((self.height, self.width), dtype=np.uint32)
        _check(self.ctx.lib.glimp_framebuffer_read_layer(self.handle, layer, _ptr(out, u32p), int(linear),
                                                         _stream_ptr(stream)), "glimp_framebuffer_read_layer")
        return out

    def resolve(self, scene: Scene, object_to_clip, mode: int = GLIMP_RESOLVE_ALBEDO, lighting=None, stream=None,
                skybox: "Texture | None" = None):
        """ShadingContext::Resolve (lit) or ResolveDebug (the other modes);
        ``lighting`` supplies object_to_world for the Normals and
        MetallicRoughness layers as well."""
        d, keep = resolve_desc(object_to_clip, mode, lighting, skybox)
        _check(self.ctx.lib.glimp_resolve(self.ctx.handle, self.handle, scene.handle, C.byref(d), _stream_ptr(stream)),
               "glimp_resolve")

    def read_color(self, stream=None) -> np.ndarray:
        out = np.zeros((self.height, self.width), dtype=np.uint32)
        _check(self.ctx.lib.glimp_framebuffer_read_color(self.handle, _ptr(out, u32p), _stream_ptr(stream)),
               "glimp_framebuffer_read_color")
        return out

    def color_ptr(self) -> int:
        p = C.c_void_p()
        _check(self.ctx.lib.glimp_framebuffer_color(self.handle, C.byref(p)), "glimp_framebuffer_color")
        return p.value

    def keys_ptr(self):
        p = C.c_void_p()
        n = C.c_size_t()
        _check(self.ctx.lib.glimp_framebuffer_keys(self.handle, C.byref(p), C.byref(n)), "glimp_framebuffer_keys")
        return p.value, n.value


def render_frame(ctx: Context, dscene: Scene, view, fb: Framebuffer | None = None, cull: bool = True,
                 flags: int = 0, stream=None) -> Framebuffer:
    """Clear, cull and draw one view: the reference frame sequence
    (Main.cpp:213-240) on the device."""
    fb = fb or Framebuffer(ctx, view.width, view.height)
    fb.clear(0, 0.0, stream)
    planes = view.frustum_planes() if cull else None
    fb.draw(dscene, view.object_to_clip, view.meshlet_count, view.meshlet_offset, planes, flags, stream)
    return fb


class Batch:
    """Many views rendered by one launch per pipeline stage, replayable as a
    single hipGraph (glimp_batch_* in include/glimp.h). ``views`` is a list
    of ``(scene, framebuffer, view)`` or ``(scene, framebuffer, view, planes,
    flags)``; every view needs its own framebuffer, and scenes and
    framebuffers must outlive the batch."""

    def __init__(self, ctx: Context, views, shape: dict | None = None, clear_color: int = 0,
                 clear_depth: float = 0.0):
        self.ctx = ctx
        n = len(views)
        arr = (ViewDesc * max(1, n))()
        self._keep = []
        for i, v in enumerate(views):
            sc, fb, vw = v[0], v[1], v[2]
            planes = v[3] if len(v) > 3 else vw.frustum_planes()
            flags = v[4] if len(v) > 4 else 0
            occlusion = v[5] if len(v) > 5 else None
            d = arr[i]
            d.scene, d.framebuffer = sc.handle, fb.handle
            d.meshlet_offset, d.meshlet_count = vw.meshlet_offset, vw.meshlet_count
            for k in range(16):
                d.object_to_clip[k] = float(vw.object_to_clip[k])
            if planes is not None:
                pl = np.ascontiguousarray(planes, dtype=np.float32)
                self._keep.append(pl)
                d.frustum_planes = _ptr(pl, f32p)
            d.flags = flags
            d.clear_color, d.clear_depth = clear_color, clear_depth
            if occlusion is not None:
                d.occlusion = C.pointer(occlusion.desc)
                self._keep.append(occlusion)
        sh = None
        if shape is not None:
            sh = BatchShape(shape["bin_size"], shape["bin_capacity"], shape.get("raster_threads", 256))
        h = C.c_void_p()
        _check(ctx.lib.glimp_batch_create(ctx.handle, arr, n, C.byref(sh) if sh else None, C.byref(h)),
               "glimp_batch_create")
        self.handle = h
        self.framebuffers = [v[1] for v in views]

    def close(self):
        if self.handle:
            self.ctx.lib.glimp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, bin_size: int, bin_capacity: int, raster_threads: int = 256):
        sh = BatchShape(bin_size, bin_capacity, raster_threads)
        _check(self.ctx.lib.glimp_batch_set_shape(self.handle, C.byref(sh)), "glimp_batch_set_shape")

    def launch(self, graph: bool = True, stream=None):
        _check(self.ctx.lib.glimp_batch_launch(self.handle, int(graph), _stream_ptr(stream)), "glimp_batch_launch")

    def set_resolve(self, resolves):
        """Resolve every view after its draw, inside the batch's launch and
        graph: ``resolves`` holds one ``(object_to_clip, mode, lighting)`` per
        view, or is None to remove them."""
        if resolves is None:
            _check(self.ctx.lib.glimp_batch_set_resolve(self.handle, None), "glimp_batch_set_resolve")
            return
        arr = (ResolveDesc * len(resolves))()
        keep = []
        for i, (o2c, mode, lighting) in enumerate(resolves):
            d, k = resolve_desc(o2c, mode, lighting)
            arr[i] = d
            keep.append(k)
        _check(self.ctx.lib.glimp_batch_set_resolve(self.handle, arr), "glimp_batch_set_resolve")

    def launch_per_view(self, stream=None):
        _check(self.ctx.lib.glimp_batch_launch_per_view(self.handle, _stream_ptr(stream)),
               "glimp_batch_launch_per_view")

    def autotune(self, stream=None) -> tuple[dict, float]:
        sh = BatchShape()
        ms = C.c_float()
        _check(self.ctx.lib.glimp_batch_autotune(self.handle, _stream_ptr(stream), C.byref(sh), C.byref(ms)),
               "glimp_batch_autotune")
        return sh.as_dict(), float(ms.value)
