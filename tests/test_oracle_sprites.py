"""Pins the oracle's light sprites (Resolve, Shading.cpp:689-727) against a
float64 statement of the same drawing.

The sprite's placement and per-tile arithmetic are pinned bit for bit
against the reference's compiled form in test_oracle_fastmath.py. Here the
whole pass is checked on a frame with a point light among the patches:
outside the sprite the image equals the resolve without sprites; inside, a
pixel behind the light (smaller reverse-Z depth) in a tile that has such a
pixel within the disc is the source-over blend of the light colour with
alpha 255 (1 - a^2), a = 1 + d / r^2, d the squared distance of the pixel
centre to the sprite centre less r^2; pixels in front of the light keep
their colour.
"""
import numpy as np

import oracle
from glimpsw_amd import scene as S


def _render(sc, view, lighting):
    ofb, _, b = oracle.render_frame(sc, view)
    depth = ofb.depth()
    rgba = oracle.resolve_lit(ofb, b, view.object_to_clip, lighting)
    return rgba, depth


def _sprite_geometry(view, pos):
    W, H = view.width, view.height
    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    c = M @ np.array([*pos, 1.0])
    ldepth = c[2] / c[3]
    radius = (max(W, H) / 30.0) / c[3]
    cx = (c[0] / c[3] * 0.5 + 0.5) * W
    cy = (c[1] / c[3] * 0.5 + 0.5) * H
    ys, xs = np.mgrid[0:H, 0:W]
    d = (xs + 0.5 - cx) ** 2 + (ys + 0.5 - cy) ** 2 - radius * radius
    inside = bool(np.abs(c[:3]).max() <= c[3])
    return inside, ldepth, radius, cx, cy, d


def _light_positions(sc, view, depth):
    """One light in front of every surface near it and one partly hidden
    behind a patch (a patch corner seen through, pushed 30% farther)."""
    yield (0.4, 0.2, -3.0)
    corners = sc.positions[:, :, [0, 7, 56, 63]].transpose(0, 2, 1).reshape(-1, 3).astype(np.float64)
    for ctr in corners[np.argsort(-corners[:, 2])]:
        if not (-8.0 < ctr[2] < -1.5):
            continue
        pos = tuple(ctr * 1.3)
        inside, ld, r, cx, cy, d = _sprite_geometry(view, pos)
        if not inside or r < 3:
            continue
        disc = d < 0
        if (disc & (ld > depth)).sum() > 10 and (disc & (ld <= depth)).sum() > 10:
            yield pos
            return
    raise AssertionError("no partly hidden light position found")


def test_point_light_sprite_against_float64():
    sc, view = S.checker_scene(width=1280, height=960)
    lighting = S.default_lighting(view)
    ofb, _, b = oracle.render_frame(sc, view)
    positions = list(_light_positions(sc, view, ofb.depth()))
    W, H = view.width, view.height
    for pos in positions:
        light = lighting.lights[1]
        assert light.type == 1
        light.position = pos
        lighting.light_sprites = True
        with_sprites, depth = _render(sc, view, lighting)
        lighting.light_sprites = False
        plain, _ = _render(sc, view, lighting)
        inside, ldepth, radius, cx, cy, d = _sprite_geometry(view, pos)
        assert inside
        x0, y0 = max(int(cx - radius), 0) & ~3, max(int(cy - radius), 0) & ~3
        x1, y1 = min(int(cx + radius), W), min(int(cy + radius), H)
        ys, xs = np.mgrid[0:H, 0:W]
        front = ldepth > depth
        in_range = (xs >= x0) & (ys >= y0) & ((xs & ~3) < x1) & ((ys & ~3) < y1)
        # tile-wide condition: a pixel of the tile inside the disc and behind the light
        hit = (d < 0) & front & in_range
        tile_hit = hit.reshape(H // 4, 4, W // 4, 4).any(axis=(1, 3))
        tile_hit = np.repeat(np.repeat(tile_hit, 4, axis=0), 4, axis=1)
        blended = tile_hit & front & in_range
        assert blended.sum() > 50, pos
        # untouched pixels (outside the range, hidden tiles, in front of the light) are the plain resolve
        assert (with_sprites[~blended] == plain[~blended]).all(), pos
        # blended pixels: source over with the float64 alpha
        a = 1.0 + d / (radius * radius)
        alpha = np.clip(np.rint(255.0 * (1.0 - a * a)), 0, 255)
        col = np.clip(np.rint(np.array(light.color) * 255.0), 0, 255)
        for k in range(3):
            bg = ((plain >> (8 * k)) & 255).astype(np.float64)
            got = ((with_sprites >> (8 * k)) & 255).astype(np.float64)
            want = bg + (col[k] - bg) * alpha / 255.0
            err = np.abs(got - want)[blended]
            assert err.max() <= 1.5, (pos, k, err.max())
        centre = blended & (d < -0.9 * radius * radius)
        assert centre.any()
        for k in range(3):
            got = (with_sprites[centre] >> (8 * k)) & 255
            assert np.abs(got.astype(int) - int(col[k])).max() <= 3
    assert len(positions) == 2
