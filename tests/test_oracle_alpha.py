"""Pins the oracle's alpha-tested fragment path (FS_EncodeSurfaceId<true>,
Shading.cpp:309-331) against what the mathematics fixes.

The expressions themselves (perspective barycentrics, clipped remap, UV
lerp, implicit mip level) are pinned bit for bit against the reference's
compiled form in test_oracle_fastmath.py; here the whole draw is checked:
 - a cutoff of 0 passes every texel, so the frame equals the frame without
   the alpha test;
 - on a perspective triangle magnifying a cutout texture, a pixel survives
   exactly when the float64 bilinear alpha at its float64
   perspective-correct UV reaches the cutoff (pixels within 1.5 alpha units
   of the cutoff are left out: the reference lerps in 8.8 fixed point);
 - the alpha test only removes fragments: every pixel of the alpha frame
   shows either the surface of the plain frame or one that lies behind it.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T


def _frame(sc, view, clipping=True):
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(view.width, view.height)
    fb.clear(0, 0.0)
    bitmap, _ = oracle.cull(b, view.frustum_planes(), 0, sc.count)
    oracle.draw(fb, b, view.object_to_clip, sc.count, 0, bitmap, clipping=clipping)
    return fb.ids(), fb.depth()


@pytest.mark.parametrize("near_frac", [0.0, 0.25])
def test_cutoff_zero_is_no_alpha_test(near_frac):
    sc = T.alpha_scene(256, seed_offset=1, spread=6.0, near_frac=near_frac, cutoffs=(0, 0))
    view = S.make_view(320, 240, count=sc.count)
    ids_a, d_a = _frame(sc, view)
    plain = dataclasses.replace(sc, materials=[(0, False, 255), (1, True, 255), (1, False, 255)])
    ids_p, d_p = _frame(plain, view)
    assert (d_p > 0).sum() > 5000
    assert (ids_a == ids_p).all() and (d_a.view(np.uint32) == d_p.view(np.uint32)).all()


@pytest.mark.parametrize("near_frac", [0.0, 0.25])
def test_alpha_only_removes_fragments(near_frac):
    sc = T.alpha_scene(256, seed_offset=2, spread=6.0, near_frac=near_frac)
    view = S.make_view(320, 240, count=sc.count)
    ids_a, d_a = _frame(sc, view)
    plain = dataclasses.replace(sc, materials=[(0, False, 255), (1, True, 255), (1, False, 255)])
    ids_p, d_p = _frame(plain, view)
    changed = ids_a != ids_p
    assert changed.sum() > 500                      # the cutout carved holes
    # a changed pixel shows a farther surface (or the background)
    assert (d_a[changed] < d_p[changed]).all()
    # surfaces that vanished from a pixel are alpha-tested ones
    mid = sc.material_id[(ids_p[changed] >> 7).astype(np.int64)]
    assert (mid != 0).all()


def _bilinear_alpha(tex, U, V):
    """Float64 SampleLevel at mip 0 with linear filtering (Texture.h:506-556):
    texel centres at +1/2, the left/top neighbour clamped at 0, no right or
    bottom neighbour past the last texel."""
    W, H = tex.width, tex.height
    a0 = np.zeros((H, W))
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    a0[ys, xs] = tex.data[textures.tiled_offset(xs, ys, tex.row_shift)] >> 24
    sx = np.clip(np.nan_to_num(U * W - 0.5), 0.0, W - 1.0)   # (clamped above for extrapolated pixels only)
    sy = np.clip(np.nan_to_num(V * H - 0.5), 0.0, H - 1.0)
    tx = np.floor(sx).astype(np.int64)
    ty = np.floor(sy).astype(np.int64)
    fx = np.where(tx + 1 < W, sx - tx, 0.0)
    fy = np.where(ty + 1 < H, sy - ty, 0.0)
    tx1 = np.minimum(tx + 1, W - 1)
    ty1 = np.minimum(ty + 1, H - 1)
    top = a0[ty, tx] * (1 - fx) + a0[ty, tx1] * fx
    bot = a0[ty1, tx] * (1 - fx) + a0[ty1, tx1] * fx
    return top * (1 - fy) + bot * fy


@pytest.mark.parametrize("cutoff", [64, 128, 200])
def test_perspective_triangle_against_float64(cutoff):
    W, H = 320, 240
    view = S.make_view(W, H)
    # a slanted triangle in view space (w = -z from 1.5 to 6: strong perspective)
    P = np.array([[-1.2, -0.8, -1.5], [1.6, -0.5, -6.0], [-0.3, 1.4, -3.0]])
    uv = np.array([[0.05, 0.05], [0.95, 0.10], [0.20, 0.95]])
    tex = textures.cutout_texture(64, 4242, period=16)
    sc = S.build_patches(np.zeros((1, 3)), np.eye(3)[None], np.ones(1), np.zeros(1, np.uint32),
                         [(0, True, cutoff)], [tex])
    pos = np.zeros((1, 3, 64), np.float32)
    pos[0, :, :3] = P.T
    tc = np.zeros((1, 64), np.uint32)
    for k in range(3):
        hb = S._half_bits(uv[k])
        tc[0, k] = int(hb[0]) | (int(hb[1]) << 16)
    idx = np.zeros((1, 3, 128), np.uint8)
    idx[0, :, 0] = [0, 1, 2]
    sc = dataclasses.replace(sc, positions=pos, texcoords=tc, indices=idx, num_triangles=np.ones(1, np.uint8),
                             num_vertices=np.full(1, 3, np.uint8),
                             bounds=np.array([[0, 0, -3, 100]], np.float32))
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(W, H)
    fb.clear(0, 0.0)
    oracle.draw(fb, b, view.object_to_clip, 1, 0, None)
    hit = fb.depth() > 0

    # float64: snapped screen vertices, exact edge functions at pixel centres
    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    clip = (M @ np.concatenate([P, np.ones((3, 1))], axis=1).T).T
    w = clip[:, 3]
    hw, hh = W // 2, H // 2
    X = np.rint((clip[:, 0] / w) * hw * 16)
    Y = np.rint((clip[:, 1] / w) * hh * 16)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    sx = (px - hw) * 16 + 8.0
    sy = (py - hh) * 16 + 8.0
    det = (X[2] - X[0]) * (Y[1] - Y[0]) - (X[1] - X[0]) * (Y[2] - Y[0])

    def edge(i, j):   # opposite vertex k: proportional to its barycentric
        return ((X[j] - X[i]) * (sy - Y[i]) - (Y[j] - Y[i]) * (sx - X[i])) / -det

    b0, b1, b2 = edge(1, 2), edge(2, 0), edge(0, 1)
    inside = (b0 > 1e-9) & (b1 > 1e-9) & (b2 > 1e-9)
    pw = np.stack([b0 / w[0], b1 / w[1], b2 / w[2]])
    pw /= pw.sum(axis=0)
    U = pw[0] * uv[0, 0] + pw[1] * uv[1, 0] + pw[2] * uv[2, 0]
    V = pw[0] * uv[0, 1] + pw[1] * uv[1, 1] + pw[2] * uv[2, 1]
    alpha = _bilinear_alpha(tex, U, V)
    clear = inside & (np.abs(alpha - cutoff) > 1.5)
    assert inside.sum() > 8000 and clear.sum() > 0.9 * inside.sum()
    expect = alpha >= cutoff
    bad = clear & (hit != expect)
    assert (hit & ~inside & (np.abs(b0) > 0.01) & (np.abs(b1) > 0.01) & (np.abs(b2) > 0.01)).sum() == 0
    assert bad.sum() == 0, f"{bad.sum()} pixels disagree with the float64 alpha test"
    assert (expect & clear).sum() > 1000 and (~expect & clear).sum() > 1000
