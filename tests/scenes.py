"""Scene variants for parity tests (built from the seeded generator)."""
from __future__ import annotations

import dataclasses

import numpy as np

from glimpsw_amd import scene as S
from glimpsw_amd import textures


def small_scene(n: int = 512, seed_offset: int = 0, spread: float = 10.0, near_frac: float = 0.0,
                double_sided: bool = False):
    """Checker-scene patches; `near_frac` of them straddle the camera plane
    beside the viewer so the guard-band clipper gets work."""
    rng = S.SplitMix64(S.SEED_BASE + 100 + seed_offset)
    centers = rng.uniform(n * 3, -spread, spread).reshape(n, 3)
    R = S._random_rotations(rng, n)
    scales = rng.uniform(n, 0.5, 2.0)
    k = int(n * near_frac)
    if k:
        s = scales[:k]
        side = np.where(rng.uniform(k) < 0.5, -1.0, 1.0)
        centers[:k, 0] = side * rng.uniform(k, 0.2, 0.9) * s
        centers[:k, 1] = rng.uniform(k, -0.5, 0.5) * s
        centers[:k, 2] = rng.uniform(k, -0.4, 0.4) * s - S.NEAR_Z
    tex = textures.checker_texture(256)
    return S.build_patches(centers, R, scales, np.zeros(n, np.uint32), [(0, double_sided, 255)], [tex])


def ragged(scene, seed: int = 7):
    """Random triangle counts (1..98) and vertex counts (>= highest index)."""
    rng = S.SplitMix64(S.SEED_BASE + 500 + seed)
    n = scene.count
    nt = (1 + (rng.next_u64(n) % np.uint64(98))).astype(np.uint8)
    idx = scene.indices.copy()
    nv = np.zeros(n, dtype=np.uint8)
    for i in range(n):
        idx[i, :, nt[i]:] = 0
        nv[i] = int(idx[i, :, : nt[i]].max()) + 1
    return dataclasses.replace(scene, num_triangles=nt, num_vertices=nv, indices=np.ascontiguousarray(idx))


def fullscreen_grid(width: int, height: int, cells: int = 7, jitter_seed: int = 3):
    """A single plane facing the camera, tessellated into meshlet patches that
    share vertices along their borders, covering the whole viewport: every
    pixel must be covered exactly once (watertightness)."""
    rng = S.SplitMix64(S.SEED_BASE + 900 + jitter_seed)
    aspect = width / height
    f = 1.0 / np.tan(np.radians(30.0))
    depth = 3.0
    half_w = depth * aspect / f * 1.3
    half_h = depth / f * 1.3
    nx = ny = cells
    # shared grid of (7*nx+1) x (7*ny+1) vertices with jittered interior
    gx = np.linspace(-half_w, half_w, 7 * nx + 1)
    gy = np.linspace(-half_h, half_h, 7 * ny + 1)
    X, Y = np.meshgrid(gx, gy)
    jit = rng.uniform(X.size * 2, -0.3, 0.3).reshape(2, *X.shape)
    dx, dy = gx[1] - gx[0], gy[1] - gy[0]
    interior = np.zeros_like(X, dtype=bool)
    interior[1:-1, 1:-1] = True
    X = X + np.where(interior, jit[0] * dx, 0)
    Y = Y + np.where(interior, jit[1] * dy, 0)
    Z = -depth + 0.2 * np.sin(X) * np.cos(Y)
    m = nx * ny
    positions = np.zeros((m, 3, 64), dtype=np.float32)
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            sl = (slice(7 * j, 7 * j + 8), slice(7 * i, 7 * i + 8))
            positions[k, 0] = X[sl].reshape(64)
            positions[k, 1] = Y[sl].reshape(64)
            positions[k, 2] = Z[sl].reshape(64)
    sc = S.build_patches(np.zeros((m, 3)), np.tile(np.eye(3), (m, 1, 1)), np.ones(m),
                         np.zeros(m, np.uint32), [(0, True, 255)], [textures.checker_texture(64)])
    bounds = np.zeros((m, 4), dtype=np.float32)
    bounds[:, :3] = positions.mean(axis=2)
    bounds[:, 3] = 1e6
    return dataclasses.replace(sc, positions=positions, bounds=bounds)


def alpha_scene(n: int = 512, seed_offset: int = 0, spread: float = 10.0, near_frac: float = 0.0,
                cutoffs=(128, 200), tex_size: int = 128):
    """small_scene with alpha-tested materials: patch i uses material i % 3,
    0 opaque (checker), 1 double-sided cutout with cutoffs[0], 2 single-sided
    cutout with cutoffs[1] (a cutoff of 255 turns the alpha test off)."""
    sc = small_scene(n, seed_offset, spread, near_frac)
    mats = [(0, False, 255), (1, True, cutoffs[0]), (1, False, cutoffs[1])]
    texs = [sc.textures[0], textures.cutout_texture(tex_size, S.SEED_BASE + 700 + seed_offset)]
    mid = (np.arange(n) % 3).astype(np.uint32)
    return dataclasses.replace(sc, material_id=mid, materials=mats, textures=texs)


def normal_mapped_scene(n: int = 384, seed_offset: int = 0, spread: float = 6.0, near_frac: float = 0.0,
                        mixed: bool = False):
    """small_scene whose material texture has a normal/metallic/roughness
    layer, with seeded tangent handedness bits; `mixed` alternates it with
    the one-layer checker so tiles see both kinds."""
    sc = small_scene(n, seed_offset, spread, near_frac)
    rng = S.SplitMix64(S.SEED_BASE + 1300 + seed_offset)
    nmr = textures.nmr_texture(128, S.SEED_BASE + 1400 + seed_offset)
    if mixed:
        mats, texs = [(0, False, 255), (1, False, 255)], [nmr, sc.textures[0]]
        mid = (np.arange(n) % 2).astype(np.uint32)
    else:
        mats, texs, mid = [(0, False, 255)], [nmr], np.zeros(n, np.uint32)
    return dataclasses.replace(sc, materials=mats, textures=texs, material_id=mid,
                               tangent_handedness=rng.next_u64(n).astype(np.uint64))


def occluder_scene(n: int = 600, seed_offset: int = 0):
    """Patches spread in front of a camera at the origin (depth 3 to 20,
    inside a widened frustum) behind a large double-sided wall at depth 2.5
    that covers the left half of the view: occlusion culling has meshlets
    to remove and meshlets that stay."""
    rng = S.SplitMix64(S.SEED_BASE + 1500 + seed_offset)
    depth = rng.uniform(n, 3.0, 20.0)
    ndc = rng.uniform(2 * n, -1.1, 1.1).reshape(n, 2)
    f = 1.0 / np.tan(np.radians(30.0))
    centers = np.stack([ndc[:, 0] * depth * (4 / 3) / f, ndc[:, 1] * depth / f, -depth], axis=1)
    R = S._random_rotations(rng, n)
    scales = rng.uniform(n, 0.3, 1.5)
    wall_c = np.array([[-1.2, 0.0, -2.5]])
    centers = np.concatenate([centers, wall_c])
    R = np.concatenate([R, np.eye(3)[None]])
    scales = np.concatenate([scales, [2.6]])
    mats = [(0, False, 255), (0, True, 255)]
    mid = np.zeros(n + 1, np.uint32)
    mid[-1] = 1
    return S.build_patches(centers, R, scales, mid, mats, [textures.checker_texture(256)])


def torus_mesh(R: float = 1.0, r: float = 0.35, nu: int = 96, nv: int = 48):
    """An indexed torus around the y axis: positions, unit normals, tangents
    (xyz along u, w = +1 on one half of the ring and -1 on the other), UVs
    and a triangle list (two triangles per grid cell, shared vertices)."""
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, V = np.meshgrid(u, v, indexing="ij")
    cx, cz = np.cos(U), np.sin(U)
    pos = np.stack([(R + r * np.cos(V)) * cx, r * np.sin(V), (R + r * np.cos(V)) * cz], axis=-1).reshape(-1, 3)
    nrm = np.stack([np.cos(V) * cx, np.sin(V), np.cos(V) * cz], axis=-1).reshape(-1, 3)
    tan = np.stack([-np.sin(U), np.zeros_like(U), np.cos(U), np.where(U < np.pi, 1.0, -1.0)], axis=-1).reshape(-1, 4)
    uv = np.stack([U / (2 * np.pi) * 4, V / (2 * np.pi) * 2], axis=-1).reshape(-1, 2)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    tris = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return (pos.astype(np.float32), nrm.astype(np.float32), tan.astype(np.float32), uv.astype(np.float32),
            tris.astype(np.uint32))
