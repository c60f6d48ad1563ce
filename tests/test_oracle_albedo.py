"""Pins the oracle's vis-buffer albedo resolve end to end (ResolveDebug
BaseColor, Shading.cpp:471-579, 749-750: IntersectTriangle's perspective
barycentrics and their screen derivatives, the UV interpolation,
CalcMipLevel, the SurfaceSampler's filter choice and SampleLevel,
Texture.h:271-280, 412-459) against float64 geometry.

 - Magnified: a slanted perspective triangle over a 64x64 linear ramp
   texture (red = 4x, green = 4y). Every covered pixel must hold the float64
   bilinear value at its float64 perspective-correct UV (texel centres at
   +0.5, level 0) within 1.5 units.
 - Minified: a far, slanted triangle over a 256x256 texture whose mip levels
   are filled with distinct constant colours (mip levels are inputs of the
   texture, so any content is allowed). Every covered pixel must show the
   colour of level clamp(floor(log2(footprint)), 0, levels - 1), except
   pixels within 0.01 of a level boundary. The footprint is the longer of
   the two float64 UV steps in texels to the neighbouring pixel centres:
   IntersectTriangle (Shading.cpp:417-464) differentiates its barycentrics
   as the difference to the pixel one to the right (ddx, screen step +2/W)
   and to the pixel one row up (ddy, step -2/H). Dropping CalcMipLevel's
   halving of ilog2, or steps of two pixels, moves whole bands to the wrong
   level.
"""
import dataclasses

import numpy as np

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures


def _triangle_scene(P, uv, tex):
    sc = S.build_patches(np.zeros((1, 3)), np.eye(3)[None], np.ones(1), np.zeros(1, np.uint32), [(0, True, 255)],
                         [tex])
    pos = np.zeros((1, 3, 64), np.float32)
    pos[0, :, :3] = np.asarray(P).T
    tc = np.zeros((1, 64), np.uint32)
    for k in range(3):
        hb = S._half_bits(np.asarray(uv[k], np.float64))
        tc[0, k] = int(hb[0]) | (int(hb[1]) << 16)
    idx = np.zeros((1, 3, 128), np.uint8)
    idx[0, :, 0] = [0, 1, 2]
    return dataclasses.replace(sc, positions=pos, texcoords=tc, indices=idx, num_triangles=np.ones(1, np.uint8),
                               num_vertices=np.full(1, 3, np.uint8), bounds=np.array([[0, 0, -3, 1000]], np.float32))


def _albedo(sc, view):
    fb, _, b = oracle.render_frame(sc, view, cull_first=False)
    hit = fb.depth() > 0
    return oracle.resolve_albedo(fb, b, view.object_to_clip), hit


def _perspective_uv(P, uv, view, dx=0.0, dy=0.0):
    """Float64 perspective-correct UV at every pixel centre (+ dx, dy pixels)
    of the triangle P, from its unsnapped screen projection, and the affine
    screen barycentrics (for inside tests)."""
    W, H = view.width, view.height
    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    clip = (M @ np.concatenate([np.asarray(P, np.float64), np.ones((3, 1))], axis=1).T).T
    w = clip[:, 3]
    sx, sy = clip[:, 0] / w, clip[:, 1] / w
    px, py = np.meshgrid(np.arange(W) + 0.5 + dx, np.arange(H) + 0.5 + dy)
    nx, ny = px * 2.0 / W - 1.0, py * 2.0 / H - 1.0
    det = (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0])
    l1 = ((nx - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (ny - sy[0])) / det
    l2 = ((sx[1] - sx[0]) * (ny - sy[0]) - (nx - sx[0]) * (sy[1] - sy[0])) / det
    lam = np.stack([1 - l1 - l2, l1, l2])
    pw = lam / w[:, None, None]
    b = pw / pw.sum(axis=0)
    uv = np.asarray(uv, np.float64)
    U = (b * uv[:, 0, None, None]).sum(axis=0)
    V = (b * uv[:, 1, None, None]).sum(axis=0)
    return U, V, lam


def test_magnified_bilinear_against_float64():
    W, H = 320, 240
    view = S.make_view(W, H, count=1)
    P = np.array([[-1.2, -0.8, -1.5], [1.6, -0.5, -6.0], [-0.3, 1.4, -3.0]])
    uv = np.array([[0.05, 0.05], [0.95, 0.10], [0.20, 0.95]])
    y, x = np.meshgrid(np.arange(64, dtype=np.uint32), np.arange(64, dtype=np.uint32), indexing="ij")
    tex = textures.create([(4 * x) | ((4 * y) << 8) | (128 << 16) | (255 << 24)])
    col, hit = _albedo(_triangle_scene(P, uv, tex), view)
    U, V, lam = _perspective_uv(P, uv, view)
    inside = (lam > 0.01).all(axis=0)
    assert inside.sum() > 8000 and (hit | ~inside).all()
    want_r = np.clip(4 * (U * 64 - 0.5), 0, 252)
    want_g = np.clip(4 * (V * 64 - 0.5), 0, 252)
    r = (col & 255).astype(np.float64)
    g = ((col >> 8) & 255).astype(np.float64)
    assert np.abs(r - want_r)[inside].max() <= 1.5
    assert np.abs(g - want_g)[inside].max() <= 1.5
    assert ((col[inside] >> 16) & 255 == 128).all()


def _level_texture(size=256):
    tex = textures.create([np.zeros((size, size), np.uint32)])
    data = tex.data.copy()
    markers = []
    for m in range(tex.mip_levels):
        # the level's whole span: TiledY8 rows of levels below 8 texels high
        # leave gaps in their 64-texel aligned region
        end = int(tex.mip_offsets[m + 1]) if m + 1 < tex.mip_levels else tex.layer_stride
        c = (10 + 35 * m) | ((250 - 30 * m) << 8) | ((40 * m) << 16) | (255 << 24)
        data[int(tex.mip_offsets[m]):end] = c
        markers.append(c)
    return dataclasses.replace(tex, data=data), markers


def _minified_case(step=1.0, halve=True):
    W, H = 320, 240
    view = S.make_view(W, H, count=1)
    P = np.array([[-3.0, -2.0, -6.0], [9.0, -2.5, -60.0], [-4.0, 3.0, -25.0]])
    uv = np.array([[0.0, 0.0], [6.0, 0.5], [0.5, 5.0]])
    tex, markers = _level_texture(256)
    col, hit = _albedo(_triangle_scene(P, uv, tex), view)
    U, V, lam = _perspective_uv(P, uv, view)
    Ux, Vx, _ = _perspective_uv(P, uv, view, dx=step)
    Uy, Vy, _ = _perspective_uv(P, uv, view, dy=-step)
    gx = np.hypot((Ux - U) * 256, (Vx - V) * 256)
    gy = np.hypot((Uy - U) * 256, (Vy - V) * 256)
    lv = np.log2(np.maximum(gx, gy)) * (1.0 if halve else 2.0)
    return tex, markers, col, hit, lam, lv


def test_minified_mip_levels_against_float64():
    tex, markers, col, hit, lam, lv = _minified_case()
    inside = (lam > 0.01).all(axis=0) & hit
    clean = inside & (np.abs(lv - np.rint(lv)) > 0.01)
    level = np.clip(np.floor(lv), 0, tex.mip_levels - 1).astype(np.int64)
    want = np.array(markers, np.uint32)[level]
    seen = np.unique(level[clean])
    assert len(seen) >= 4, seen
    bad = clean & (col != want)
    assert not bad.any(), f"{bad.sum()} of {clean.sum()} pixels off their level"
    assert clean.sum() > 0.95 * inside.sum()


def test_mip_level_check_discriminates():
    """Models with two-pixel steps or without the halving of ilog2 put a
    large share of the pixels on the wrong level."""
    for kw in ({"step": 2.0}, {"halve": False}):
        tex, markers, col, hit, lam, lv = _minified_case(**kw)
        inside = (lam > 0.01).all(axis=0) & hit
        level = np.clip(np.floor(lv), 0, tex.mip_levels - 1).astype(np.int64)
        want = np.array(markers, np.uint32)[level]
        assert (inside & (col != want)).sum() > 0.3 * inside.sum(), kw
