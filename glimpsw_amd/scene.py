"""Seeded synthetic scenes for the meshlet raster path.

A scene is a structure-of-arrays copy of the reference's ``Meshlet`` record
(Scene.h:15-30): per meshlet a bound sphere, vertex/triangle counts, a
material id, ``Positions[3][64]`` fp32, ``TexCoords[64]`` half2,
``NormalTangents[64]`` (octahedral normal and tangent, 8 bits each) and
``Indices[3][128]`` u8. Every patch is an 8x8-vertex grid (64 vertices,
98 triangles) with UVs equal to the grid coordinates.

Randomness comes from splitmix64 with ``seed = 0x474C494D0000 + config`` and
uniform floats ``(x >> 40) * 2**-24``; only IEEE arithmetic and square roots
are used, so a scene is bit-identical on every host.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import textures

SEED_BASE = 0x474C494D0000
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
NEAR_Z = 0.01


class SplitMix64:
    """Counter-based splitmix64 stream (the mixer of Steele et al., 2014)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            z = self.state + k * GOLDEN
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
        return lo + (hi - lo) * u

    def log_uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        # exp(U(log lo, log hi)) without libm: lo * (hi/lo)**u via float64 pow
        u = self.uniform(n)
        return lo * np.power(hi / lo, u)


@dataclasses.dataclass
class Scene:
    bounds: np.ndarray           # (M,4) f32
    num_vertices: np.ndarray     # (M,) u8
    num_triangles: np.ndarray    # (M,) u8
    material_id: np.ndarray      # (M,) u32
    tangent_handedness: np.ndarray  # (M,) u64
    positions: np.ndarray        # (M,3,64) f32
    texcoords: np.ndarray        # (M,64) u32 (half2)
    normal_tangents: np.ndarray  # (M,64) u32
    indices: np.ndarray          # (M,3,128) u8
    materials: list              # [(texture index, double_sided, alpha_cutoff)]
    textures: list               # [textures.Texture]

    @property
    def count(self) -> int:
        return int(self.bounds.shape[0])

    def triangles(self) -> int:
        return int(self.num_triangles.astype(np.int64).sum())


@dataclasses.dataclass
class View:
    width: int
    height: int
    proj: np.ndarray             # (16,) f32 column-major, reverse-Z infinite
    view: np.ndarray             # (16,) f32 column-major world->view
    object_to_clip: np.ndarray   # (16,) f32 column-major = proj * view (model = I)
    meshlet_offset: int = 0
    meshlet_count: int = 0

    def frustum_planes(self) -> np.ndarray:
        return frustum_planes(self.object_to_clip)


def proj_matrix(fov_deg: float, aspect: float, near: float = NEAR_Z) -> np.ndarray:
    """Reverse-Z infinite perspective (Camera.h:38-48), column-major."""
    f = 1.0 / math.tan(math.radians(fov_deg) / 2.0)
    m = np.zeros(16, dtype=np.float64)
    m[0] = f / aspect
    m[5] = -f
    m[11] = -1.0
    m[14] = near
    return m.astype(np.float32)


def mat_mul_cm(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major 4x4 product a*b in float64, rounded to float32."""
    A = a.astype(np.float64).reshape(4, 4).T
    B = b.astype(np.float64).reshape(4, 4).T
    return (A @ B).T.reshape(16).astype(np.float32)


def frustum_planes(obj_to_clip: np.ndarray) -> np.ndarray:
    """Gribb-Hartmann planes as in CullMeshlets (Shading.cpp:783-791):
    rows of the object->clip matrix, row3 +- row_i, normalised by |xyz|;
    order left, right, bottom, top, near (the fifth of six is unused there
    too). Computed in float64 and rounded once."""
    M = obj_to_clip.astype(np.float64).reshape(4, 4).T  # M[r, c]
    planes = []
    for i in range(3):
        a = M[3] + M[i]
        b = M[3] - M[i]
        planes.append(a / np.linalg.norm(a[:3]))
        planes.append(b / np.linalg.norm(b[:3]))
    return np.stack(planes[:5]).astype(np.float32).reshape(20)


def view_matrix(eye, yaw: float, pitch: float) -> np.ndarray:
    """World->view: rotate by yaw about Y, pitch about X, translate by -eye."""
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    Ry = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]], dtype=np.float64)
    Rx = np.array([[1, 0, 0], [0, cp, sp], [0, -sp, cp]], dtype=np.float64)
    R = Rx @ Ry
    t = -R @ np.asarray(eye, dtype=np.float64)
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    return M.T.reshape(16).astype(np.float32)


def _half_bits(x: np.ndarray) -> np.ndarray:
    return x.astype(np.float16).view(np.uint16).astype(np.uint32)


def _oct_encode(n: np.ndarray) -> np.ndarray:
    """Octahedral map of unit vectors to [0,1]^2 (Texture.h:282-288 inverse
    of UnmapOctahedron), float64."""
    s = np.abs(n).sum(axis=-1, keepdims=True)
    p = n / s
    u, v, z = p[..., 0], p[..., 1], p[..., 2]
    neg = z < 0
    su = np.where(u >= 0, 1.0, -1.0)
    sv = np.where(v >= 0, 1.0, -1.0)
    uo = np.where(neg, (1 - np.abs(v)) * su, u)
    vo = np.where(neg, (1 - np.abs(u)) * sv, v)
    return np.stack([uo * 0.5 + 0.5, vo * 0.5 + 0.5], axis=-1)


def _grid_indices() -> np.ndarray:
    idx = np.zeros((3, 128), dtype=np.uint8)
    t = 0
    for j in range(7):
        for i in range(7):
            v00 = j * 8 + i
            v10, v11, v01 = v00 + 1, v00 + 9, v00 + 8
            for tri in ((v00, v10, v11), (v00, v11, v01)):
                idx[:, t] = tri
                t += 1
    assert t == 98
    return idx


GRID_INDICES = _grid_indices()


def _random_rotations(rng: SplitMix64, n: int) -> np.ndarray:
    """Uniform random rotations from rejection-sampled quaternions (no libm
    transcendental calls). Returns (n,3,3) float64."""
    cand = 48
    q = rng.uniform(n * cand * 4, -1.0, 1.0).reshape(n, cand, 4)
    r2 = (q * q).sum(-1)
    ok = (r2 <= 1.0) & (r2 > 1e-4)
    first = np.argmax(ok, axis=1)
    assert ok[np.arange(n), first].all(), "rotation rejection sampling failed"
    q = q[np.arange(n), first]
    q = q / np.sqrt((q * q).sum(-1, keepdims=True))
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def build_patches(centers: np.ndarray, R: np.ndarray, scales: np.ndarray, material_id: np.ndarray,
                  mats: list, texs: list) -> Scene:
    """Grid patches: local ((i/7-1/2)s, (j/7-1/2)s, 0), rotated, translated."""
    n = centers.shape[0]
    g = (np.arange(8, dtype=np.float64) / 7.0) - 0.5
    gx = np.tile(g, 8)        # i = v % 8
    gy = np.repeat(g, 8)      # j = v // 8
    local = np.stack([gx, gy, np.zeros(64)], axis=0)          # (3,64)
    local = local[None] * scales[:, None, None]               # (n,3,64)
    world = np.einsum("nij,njv->niv", R, local) + centers[:, :, None]
    positions = world.astype(np.float32)

    uv = (_half_bits(np.tile(np.arange(8.0), 8)) | (_half_bits(np.repeat(np.arange(8.0), 8)) << 16)).astype(np.uint32)
    texcoords = np.broadcast_to(uv, (n, 64)).copy()

    nrm = R[:, :, 2]
    tan = R[:, :, 0]
    on = np.clip(np.rint(_oct_encode(nrm) * 255.0), 0, 255).astype(np.uint32)
    ot = np.clip(np.rint(_oct_encode(tan) * 255.0), 0, 255).astype(np.uint32)
    packed = on[:, 0] | (on[:, 1] << 8) | (ot[:, 0] << 16) | (ot[:, 1] << 24)
    normal_tangents = np.broadcast_to(packed[:, None], (n, 64)).copy().astype(np.uint32)

    radius = (scales * math.sqrt(0.5) * 1.0001).astype(np.float32)
    bounds = np.concatenate([centers.astype(np.float32), radius[:, None]], axis=1)
    return Scene(
        bounds=np.ascontiguousarray(bounds, dtype=np.float32),
        num_vertices=np.full(n, 64, dtype=np.uint8),
        num_triangles=np.full(n, 98, dtype=np.uint8),
        material_id=material_id.astype(np.uint32),
        tangent_handedness=np.zeros(n, dtype=np.uint64),
        positions=np.ascontiguousarray(positions),
        texcoords=texcoords,
        normal_tangents=normal_tangents,
        indices=np.ascontiguousarray(np.broadcast_to(GRID_INDICES, (n, 3, 128))),
        materials=mats,
        textures=texs,
    )


def make_view(width: int, height: int, fov_deg: float = 60.0, view: np.ndarray | None = None,
              offset: int = 0, count: int = 0) -> View:
    proj = proj_matrix(fov_deg, width / height)
    v = view if view is not None else np.eye(4, dtype=np.float32).reshape(16)
    return View(width, height, proj, v, mat_mul_cm(proj, v), offset, count)


def checker_scene(num_meshlets: int = 512, width: int = 640, height: int = 480) -> tuple[Scene, View]:
    """Plain pipeline, CPU-runnable: patches uniform in [-10,10]^3, random
    rotation, scale 0.5-2; camera at the origin, 60 deg FOV; one 256x256
    checker texture with its full mip chain."""
    rng = SplitMix64(SEED_BASE + 0)
    centers = rng.uniform(num_meshlets * 3, -10.0, 10.0).reshape(num_meshlets, 3)
    R = _random_rotations(rng, num_meshlets)
    scales = rng.uniform(num_meshlets, 0.5, 2.0)
    tex = textures.checker_texture(256)
    scene = build_patches(centers, R, scales, np.zeros(num_meshlets, np.uint32), [(0, False, 255)], [tex])
    return scene, make_view(width, height, count=num_meshlets)


def frame_1080p(num_meshlets: int = 10240, width: int = 1920, height: int = 1080, tex_size: int = 2048,
            num_materials: int = 16, crossers_inside: bool = False) -> tuple[Scene, View]:
    """Single 1080p view of ~1M triangles: per-patch triangle area
    log-uniform in [0.5, 500] px, patches spread over a widened frustum (so
    culling and outcode rejection see work), random orientation (about half
    back-facing) and 5% of patches straddling the camera plane to exercise
    near-plane clipping; 16 seeded noise materials with mips.

    The straddling patches sit beside the eye, their centres 0.85-1.05 of
    their size off the view axis, so most of their triangles are rejected or
    accepted without clipping and the depth complexity stays about 3. With
    `crossers_inside` they sit 0.5-0.7 of their size off the axis, reaching
    into the frustum: about ten times the clipped triangles, each covering a
    large part of the screen (depth complexity about 18); the rest of the
    scene is the same."""
    rng = SplitMix64(SEED_BASE + 1)
    n = num_meshlets
    aspect = width / height
    f = 1.0 / math.tan(math.radians(60.0) / 2.0)
    area = rng.log_uniform(n, 0.5, 500.0)               # target px per triangle (face-on)
    depth = rng.log_uniform(n, 1.0, 60.0)
    ndc = rng.uniform(2 * n, -2.0, 2.0).reshape(n, 2)
    R = _random_rotations(rng, n)
    px_per_unit = (height / 2.0) * f / depth
    scales = 7.0 * np.sqrt(2.0 * area) / px_per_unit
    centers = np.stack([ndc[:, 0] * depth * aspect / f, ndc[:, 1] * depth / f, -depth], axis=1)
    # near-plane crossers: beside the camera, half in front of it
    near_sel = rng.uniform(n) < 0.05
    k = int(near_sel.sum())
    s_near = rng.uniform(k, 0.5, 2.0)
    side = np.where(rng.uniform(k) < 0.5, -1.0, 1.0)
    off = (0.5, 0.7) if crossers_inside else (0.85, 1.05)
    cx = side * rng.uniform(k, off[0], off[1]) * s_near
    cy = rng.uniform(k, -0.4, 0.4) * s_near
    cz = rng.uniform(k, -0.3, 0.3) * s_near - NEAR_Z
    centers[near_sel] = np.stack([cx, cy, cz], axis=1)
    scales[near_sel] = s_near
    mats = [(i, False, 255) for i in range(num_materials)]
    texs = [textures.noise_texture(tex_size, SEED_BASE + 1000 + i) for i in range(num_materials)]
    mid = (np.arange(n) % num_materials).astype(np.uint32)
    scene = build_patches(centers, R, scales, mid, mats, texs)
    return scene, make_view(width, height, count=n)


def batch_views(num_scenes: int = 8, num_views: int = 64, min_meshlets: int = 1000, max_meshlets: int = 50000,
            min_res: tuple = (256, 256), max_res: tuple = (1920, 1080)) -> tuple[list, list]:
    """Batching workload: ``num_scenes`` seeded scenes with log-uniform
    meshlet counts in [min_meshlets, max_meshlets], ``num_views`` views (view
    v shows scene v % num_scenes) with per-axis log-uniform resolution
    between ``min_res`` and ``max_res`` (rounded down to multiples of 4; one
    draw per view of its whole scene) and random cameras inside the scene. Patches
    are as in checker_scene (scale 0.5-2, random rotation) at its density
    (512 per 20^3), in a cube sized for the count; every scene has one
    256x256 checker texture. Returns (scenes, [(scene index, View)])."""
    rng = SplitMix64(SEED_BASE + 2)
    counts = np.floor(rng.log_uniform(num_scenes, min_meshlets, max_meshlets)).astype(np.int64)
    tex = textures.checker_texture(256)
    scenes = []
    sides = []
    for n in counts:
        n = int(n)
        side = 20.0 * (n / 512.0) ** (1.0 / 3.0)
        centers = rng.uniform(n * 3, -side / 2, side / 2).reshape(n, 3)
        R = _random_rotations(rng, n)
        scales = rng.uniform(n, 0.5, 2.0)
        scenes.append(build_patches(centers, R, scales, np.zeros(n, np.uint32), [(0, False, 255)], [tex]))
        sides.append(side)
    views = []
    for v in range(num_views):
        si = v % num_scenes
        tw, th = rng.uniform(2)
        # per-axis log-uniform size, rounded down to the 4x4 tile
        w = max(4, 4 * int(min_res[0] * (max_res[0] / min_res[0]) ** float(tw) / 4.0))
        h = max(4, 4 * int(min_res[1] * (max_res[1] / min_res[1]) ** float(th) / 4.0))
        eye = rng.uniform(3, -0.4 * sides[si], 0.4 * sides[si])
        yaw = float(rng.uniform(1, 0.0, 2.0 * math.pi)[0])
        pitch = float(rng.uniform(1, -0.6, 0.6)[0])
        views.append((si, make_view(w, h, 60.0, view_matrix(eye, yaw, pitch), 0, scenes[si].count)))
    return scenes, views


def stress_views(num_views: int = 16, width: int = 2880, height: int = 1620, patches_per_layer: int = 5120,
                 layers: int = 8, subpixel_frac: float = 0.30, tex_size: int = 512,
                 num_materials: int = 16) -> tuple[Scene, list]:
    """Overdraw stress: ``layers`` depth-stacked layers of grid patches
    (depths 2 * 1.35^l), each layer spread over the frustum densely enough
    that its front faces cover the screen about once (depth complexity about
    ``layers``); a ``subpixel_frac`` share of patches is shrunk 12x so their
    triangles are sub-pixel. ``num_views`` views near the layout's camera
    (yaw and pitch within 3 degrees) at ``width`` x ``height``, close to the
    2896-pixel fixed-point limit. Returns (scene, [View])."""
    rng = SplitMix64(SEED_BASE + 4)
    n = patches_per_layer * layers
    aspect = width / height
    f = 1.0 / math.tan(math.radians(60.0) / 2.0)
    layer = np.repeat(np.arange(layers), patches_per_layer)
    depth = 2.0 * 1.35 ** layer.astype(np.float64)
    xy = rng.uniform(2 * n, -1.05, 1.05).reshape(n, 2)
    jitter = rng.uniform(n, -0.05, 0.05)
    centers = np.stack([xy[:, 0] * depth * aspect / f, xy[:, 1] * depth / f, -depth * (1.0 + jitter)], axis=1)
    R = _random_rotations(rng, n)
    scales = depth * 0.049 * rng.uniform(n, 0.8, 1.25)
    tiny = rng.uniform(n) < subpixel_frac
    scales[tiny] /= 12.0
    mats = [(i, False, 255) for i in range(num_materials)]
    texs = [textures.noise_texture(tex_size, SEED_BASE + 4000 + i) for i in range(num_materials)]
    mid = (np.arange(n) % num_materials).astype(np.uint32)
    scene = build_patches(centers, R, scales, mid, mats, texs)
    views = []
    for _ in range(num_views):
        yaw = float(rng.uniform(1, -0.05, 0.05)[0])
        pitch = float(rng.uniform(1, -0.05, 0.05)[0])
        views.append(make_view(width, height, 60.0, view_matrix((0.0, 0.0, 0.0), yaw, pitch), 0, n))
    return scene, views


def inverse_screen_proj(world_to_clip: np.ndarray, width: int, height: int) -> np.ndarray:
    """GetInverseScreenProjMatrix (Camera.h:140-146): inverse(M) then
    translate(-1,-1,0), scale(2/size, 1), translate(0.5, 0.5, 0), so pixel
    coordinates (x, y, depth, 1) map to world space. float64, rounded once."""
    M = world_to_clip.astype(np.float64).reshape(4, 4).T
    inv = np.linalg.inv(M)
    T1 = np.eye(4); T1[:2, 3] = -1.0
    Sc = np.eye(4); Sc[0, 0] = 2.0 / width; Sc[1, 1] = 2.0 / height
    T2 = np.eye(4); T2[:2, 3] = 0.5
    R = inv @ T1 @ Sc @ T2
    return R.T.reshape(16).astype(np.float32)


@dataclasses.dataclass
class Light:
    type: int                      # 0 directional, 1 point, 2 spot
    position: tuple = (0.0, 0.0, 0.0)
    direction: tuple = (0.0, -1.0, 0.0)
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1000.0
    radius: float = 0.0

    @property
    def inv_radius_sq(self) -> float:
        r = self.radius if self.radius > 0 else 1e6   # Light::SetRadius (Scene.h:65-68)
        return float(np.float32(1.0) / (np.float32(r) * np.float32(r)))


@dataclasses.dataclass
class Lighting:
    lights: list
    exposure: float
    view_pos: tuple
    object_to_world: np.ndarray    # (9,) mat3 column-major
    inv_screen_proj: np.ndarray    # (16,)
    world_to_clip: np.ndarray | None = None   # (16,) for the light sprites
    light_sprites: bool = True     # Resolve draws point/spot lights (Shading.cpp:689-727)
    skybox: object = None          # host R11G11B10f octahedral texture (ShadingContext::SkyboxTex), or None


def default_lighting(view: View) -> Lighting:
    """A sun and a point light near the camera; exposure 1."""
    d = np.array([-0.3, -1.0, -0.5])
    d = d / np.linalg.norm(d)
    lights = [Light(0, direction=tuple(d.astype(np.float32).tolist()), color=(1.0, 0.95, 0.9), intensity=3000.0),
              Light(1, position=(1.5, 1.0, -4.0), color=(0.6, 0.7, 1.0), intensity=4000.0, radius=30.0)]
    Vinv = np.linalg.inv(view.view.astype(np.float64).reshape(4, 4).T)
    eye = tuple(Vinv[:3, 3].astype(np.float32).tolist())
    return Lighting(lights, 1.0, eye, np.eye(3, dtype=np.float32).reshape(9),
                    inverse_screen_proj(view.object_to_clip, view.width, view.height),
                    np.asarray(view.object_to_clip, np.float32))


def occlusion_terms(proj_mat: np.ndarray, prev_view_mat: np.ndarray, model_mat: np.ndarray | None = None):
    """The terms CullMeshlets derives from its matrices (Shading.cpp:779-795),
    computed in float64 and rounded once: objectToPrevViewMat = prevView *
    model (column-major), objectToWorldScale = |model column 0| and
    projFactors = (proj[3][2], proj[0][0], proj[1][1])."""
    M = np.eye(4) if model_mat is None else model_mat.astype(np.float64).reshape(4, 4).T
    Vp = prev_view_mat.astype(np.float64).reshape(4, 4).T
    o2pv = (Vp @ M).T.reshape(16).astype(np.float32)
    scale = np.float32(np.linalg.norm(M[:3, 0]))
    P = proj_mat.astype(np.float64).reshape(16)
    proj = np.array([P[14], P[0], P[5]], np.float32)
    return o2pv, scale, proj
