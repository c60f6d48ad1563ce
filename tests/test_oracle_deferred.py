"""Pins the oracle's deferred pipeline (DeferredShader: FS_EncodeGBuffer,
Shading.cpp:344-407, 655) against what the geometry and the vis-buffer
pipeline fix.

The packing of layer 2 (MapOctahedron, 10-bit channels, 6-bit metallic and
roughness) is pinned bit for bit against the reference's compiled form in
test_oracle_fastmath.py, and so are the barycentrics, UVs, implicit mip
level and normal mapping it shares with the alpha test and the resolve.
Here whole draws are checked:
 - depth: the G-buffer's depth layer equals the vis-buffer pipeline's for
   the same alpha-tested scene (the same fragments pass), bit for bit;
 - base colour: on a slanted perspective triangle magnifying a linear
   ramp texture, every pixel holds the float64 bilinear value at its
   float64 perspective-correct UV, within 1.5 units;
 - normals: on a quad facing the camera with a constant tangent-space
   normal, layer 2 decodes to the float64 tangent-frame normal (flipped
   bitangent when the handedness bit is set) within the 10-bit encoding,
   and carries the texel's metallic and roughness bytes cut to 6 bits;
 - a meshlet without a material writes depth only: the colour layer keeps
   the colour of the surface drawn before it;
 - a clear leaves layer 2 alone, as Framebuffer::Clear clears only
   layers 0 and 1 (Rasterizer.h:35-38).
"""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T


def _deferred(sc, view, clear=0, fb=None, first=0, count=None, object_to_world=None, cull=True):
    b = oracle.SceneBinding(sc)
    if fb is None:
        fb = oracle.OracleFramebuffer(view.width, view.height, layers=3)
        fb.clear(clear, 0.0)
    count = sc.count - first if count is None else count
    bm = oracle.cull(b, view.frustum_planes(), first, count)[0] if cull else None
    ctr, _ = oracle.draw(fb, b, view.object_to_clip, count, first, bm, pipeline=oracle.PIPELINE_DEFERRED,
                         object_to_world=object_to_world)
    return fb, ctr


@pytest.mark.parametrize("near_frac", [0.0, 0.2])
def test_depth_matches_visbuffer_with_alpha(near_frac):
    sc = T.alpha_scene(384, seed_offset=2, spread=6.0, near_frac=near_frac)
    view = S.make_view(256, 192, count=sc.count)
    fb, _ = _deferred(sc, view, clear=0x11223344)
    vfb, _, _ = oracle.render_frame(sc, view)
    depth = fb.depth()
    assert (depth.view(np.uint32) == vfb.depth().view(np.uint32)).all()
    covered = depth > 0
    assert covered.sum() > 5000 and (near_frac > 0 or (~covered).sum() > 100)
    col = fb.pixels(0)
    assert (col[~covered] == 0x11223344).all()
    assert ((col[covered] >> 24) >= 128).all()      # every surface passed its alpha cutoff


def _ramp_texture():
    y, x = np.meshgrid(np.arange(64, dtype=np.uint32), np.arange(64, dtype=np.uint32), indexing="ij")
    return textures.create([(4 * x) | ((4 * y) << 8) | (128 << 16) | (255 << 24)])


def _triangle_scene(P, uv, tex, material=(0, True, 0)):
    sc = S.build_patches(np.zeros((1, 3)), np.eye(3)[None], np.ones(1), np.zeros(1, np.uint32), [material], [tex])
    pos = np.zeros((1, 3, 64), np.float32)
    pos[0, :, :3] = np.asarray(P).T
    tc = np.zeros((1, 64), np.uint32)
    for k in range(3):
        hb = S._half_bits(np.asarray(uv[k], np.float64))
        tc[0, k] = int(hb[0]) | (int(hb[1]) << 16)
    idx = np.zeros((1, 3, 128), np.uint8)
    idx[0, :, 0] = [0, 1, 2]
    return dataclasses.replace(sc, positions=pos, texcoords=tc, indices=idx, num_triangles=np.ones(1, np.uint8),
                               num_vertices=np.full(1, 3, np.uint8), bounds=np.array([[0, 0, -3, 100]], np.float32))


def test_base_colour_against_float64():
    W, H = 320, 240
    view = S.make_view(W, H)
    P = np.array([[-1.2, -0.8, -1.5], [1.6, -0.5, -6.0], [-0.3, 1.4, -3.0]])
    uv = np.array([[0.05, 0.05], [0.95, 0.10], [0.20, 0.95]])
    sc = _triangle_scene(P, uv, _ramp_texture())
    fb, _ = _deferred(sc, view, cull=False)
    hit = fb.depth() > 0
    col = fb.pixels(0)

    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    clip = (M @ np.concatenate([P, np.ones((3, 1))], axis=1).T).T
    w = clip[:, 3]
    hw, hh = W // 2, H // 2
    X = np.rint((clip[:, 0] / w) * hw * 16)
    Y = np.rint((clip[:, 1] / w) * hh * 16)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    sx, sy = (px - hw) * 16 + 8.0, (py - hh) * 16 + 8.0
    det = (X[2] - X[0]) * (Y[1] - Y[0]) - (X[1] - X[0]) * (Y[2] - Y[0])

    def edge(i, j):
        return ((X[j] - X[i]) * (sy - Y[i]) - (Y[j] - Y[i]) * (sx - X[i])) / -det

    b = np.stack([edge(1, 2), edge(2, 0), edge(0, 1)])
    inside = (b > 0.01).all(axis=0)
    pw = b / w[:, None, None]
    pw /= pw.sum(axis=0)
    U = (pw * uv[:, 0, None, None]).sum(axis=0)
    V = (pw * uv[:, 1, None, None]).sum(axis=0)
    want_r = np.clip(4 * (U * 64 - 0.5), 0, 252)
    want_g = np.clip(4 * (V * 64 - 0.5), 0, 252)
    assert inside.sum() > 8000 and (hit | ~inside).all()
    r = (col & 255).astype(np.float64)
    g = ((col >> 8) & 255).astype(np.float64)
    assert np.abs(r - want_r)[inside].max() <= 1.5
    assert np.abs(g - want_g)[inside].max() <= 1.5
    assert ((col[inside] >> 16) & 255 == 128).all() and ((col[inside] >> 24) == 255).all()


def _quad_scene(handed: bool, nmr_texel: int):
    sc = S.build_patches(np.array([[0.0, 0.0, -2.0]]), np.eye(3)[None], np.array([3.0]), np.zeros(1, np.uint32),
                         [(0, True, 0)], [])
    albedo = np.full((16, 16), 0xFFB0B0B0, np.uint32)
    tex = textures.create([albedo, np.full((16, 16), nmr_texel, np.uint32)])
    hand = np.array([np.uint64(0xFFFFFFFFFFFFFFFF) if handed else np.uint64(0)], np.uint64)
    return dataclasses.replace(sc, textures=[tex], tangent_handedness=hand)


def _decode_oct(word):
    u = (word & 1023) / 1023.0 * 2 - 1
    v = ((word >> 10) & 1023) / 1023.0 * 2 - 1
    z = 1 - np.abs(u) - np.abs(v)
    t = np.maximum(-z, 0)
    x = u - np.sign(u) * t
    y = v - np.sign(v) * t
    n = np.stack([x, y, z], axis=-1)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


@pytest.mark.parametrize("handed", [False, True])
def test_normal_layer_on_quad(handed):
    # tangent-space normal (0.5, -0.4, sqrt(1 - 0.41)) from bytes 191.25 -> 191, 76.5 -> 77
    bx, by, metal, rough = 191, 77, 0x9C, 0x37
    sc = _quad_scene(handed, bx | (by << 8) | (metal << 16) | (rough << 24))
    view = S.make_view(64, 64, count=1)
    fb, _ = _deferred(sc, view)
    assert (fb.depth() > 0).all()
    word = fb.pixels(2)
    n = _decode_oct(word)
    tx, ty = bx / 127.5 - 1, by / 127.5 - 1
    ts = np.array([tx, ty, np.sqrt(1 - tx * tx - ty * ty)])
    # the patch lies in the xy plane facing +z: N = +z, T = +x, B = cross(N, T) = +y
    B = np.array([0.0, -1.0, 0.0]) if handed else np.array([0.0, 1.0, 0.0])
    want = ts[0] * np.array([1.0, 0, 0]) + ts[1] * B + ts[2] * np.array([0, 0, 1.0])
    err = np.abs(n - want / np.linalg.norm(want)).max()
    assert err < 6e-3, err
    assert ((word >> 20) & 63 == metal >> 2).all() and ((word >> 26) == rough >> 2).all()


def test_material_less_meshlet_keeps_colour_and_layer2_survives_clear():
    view = S.make_view(64, 64, count=2)
    sc = _quad_scene(False, 0x80808080)
    far = dataclasses.replace(sc)
    # two quads: the material one far, then a material-less one nearer
    both = S.build_patches(np.array([[0.0, 0.0, -3.0], [0.0, 0.0, -2.0]]), np.stack([np.eye(3)] * 2),
                           np.array([4.0, 4.0]), np.array([0, 0xFFFFFFFF], np.uint32), [(0, True, 0)],
                           far.textures)
    fb, _ = _deferred(both, view, clear=0x01020304, cull=False)
    alone, _ = _deferred(both, view, clear=0x01020304, count=1, cull=False)
    near_only, _ = _deferred(both, view, clear=0x01020304, first=1, count=1, cull=False)
    assert (near_only.pixels(0) == 0x01020304).all() and (near_only.depth() > 0).all()
    assert (fb.pixels(0) == alone.pixels(0)).all()
    assert (fb.depth().view(np.uint32) == near_only.depth().view(np.uint32)).all()
    assert (fb.depth() > alone.depth()).all()
    before = fb.pixels(2).copy()
    assert (before != 0).any()
    fb.clear(0, 0.0)
    assert (fb.pixels(2) == before).all() and (fb.pixels(0) == 0).all()
