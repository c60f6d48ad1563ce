"""Seeded random tiny scenes for property tests: arbitrary meshlets (random
vertex and triangle counts, random indices, triangles crossing the near
plane and behind the camera, degenerate and coincident triangles), mixed
materials (single- and double-sided, alpha-tested cutouts) and random
framebuffer sizes, clear depths, draw splits and launch shapes."""
from __future__ import annotations

import dataclasses

import numpy as np

from glimpsw_amd import scene as S
from glimpsw_amd import textures


@dataclasses.dataclass
class Case:
    scene: S.Scene
    view: S.View
    draws: list           # (offset, count)
    cull: bool
    clear_depth: float
    flags: int            # GLIMP_DRAW_NO_CLIPPING / NO_GUARDBAND bits
    bin_size: int
    bin_capacity: int
    direct: bool


def make_case(seed: int) -> Case:
    rng = np.random.default_rng(10_000 + seed)
    m = int(rng.integers(1, 40))
    nv = rng.integers(3, 65, m).astype(np.uint8)
    nt = rng.integers(1, 129, m).astype(np.uint8)
    pos = np.zeros((m, 3, 64), np.float32)
    idx = np.zeros((m, 3, 128), np.uint8)
    bounds = np.zeros((m, 4), np.float32)
    for i in range(m):
        centre = rng.uniform([-1.5, -1.5, -4.0], [1.5, 1.5, 0.3])
        spread = rng.uniform(0.05, 1.5)
        p = centre[:, None] + rng.normal(0, spread, (3, 64))
        if rng.random() < 0.2:
            p[:, 32:] = p[:, :32]                                  # coincident vertices
        pos[i] = p.astype(np.float32)
        t = rng.integers(0, int(nv[i]), (3, int(nt[i])))
        if rng.random() < 0.2:
            t[1, ::5] = t[0, ::5]                                  # some degenerate triangles
        idx[i, :, : nt[i]] = t
        used = pos[i][:, : nv[i]].astype(np.float64)
        c = used.mean(axis=1)
        bounds[i, :3] = c
        bounds[i, 3] = np.float32(np.linalg.norm(used - c[:, None], axis=0).max() * 1.001 + 1e-4)
    uv = rng.uniform(-2, 4, (m, 64, 2)).astype(np.float16).view(np.uint16).astype(np.uint32)
    tc = (uv[..., 0] | (uv[..., 1] << 16)).astype(np.uint32)
    mats = [(0, False, 255), (0, True, 255), (1, False, 100), (1, True, 200)]
    texs = [textures.checker_texture(64), textures.cutout_texture(32, 777 + seed % 5)]
    mid = rng.integers(0, len(mats), m).astype(np.uint32)
    sc = S.Scene(bounds, nv, nt, mid, rng.integers(0, 2 ** 63, m).astype(np.uint64), pos, tc,
                 rng.integers(0, 2 ** 32, (m, 64), dtype=np.uint64).astype(np.uint32), idx, mats, texs)
    w = 4 * int(rng.integers(1, 65))
    h = 4 * int(rng.integers(1, 49))
    view = S.make_view(w, h, fov_deg=float(rng.uniform(30, 100)), count=m)
    cuts = sorted(set(int(x) for x in rng.integers(0, m, int(rng.integers(0, 3))))) if m > 1 else []
    edges = [0] + [c for c in cuts if 0 < c < m] + [m]
    draws = [(a, b - a) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    return Case(sc, view, draws, bool(rng.random() < 0.7),
                float(rng.choice([0.0, 0.0, 0.005, 0.05])), int(rng.choice([0, 0, 0, 1, 2])),
                int(rng.choice([8, 16, 32, 64])), int(rng.choice([2, 16, 1024])), bool(rng.random() < 0.3))
