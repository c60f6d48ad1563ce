"""Pins the oracle's normal/metallic/roughness layer in the lit resolve
(ResolveSurface, Shading.cpp:541-577) against what the geometry fixes.

The normal-mapping expression itself is pinned bit for bit against the
reference's compiled form (test_oracle_fastmath.py::test_normal_mapping).
Here a quad facing the camera (normal +z, tangent +x) carries a constant
texel whose tangent-space normal leans 45 degrees towards the bitangent,
B = cross(N, T) = +y, or -y when the vertex's handedness bit is set. A
directional light from +y+z must then light the quad almost head-on
(N.L close to 1) without the handedness bit and almost grazing (N.L close
to 0) with it, and the reverse for a light from -y+z; the texel's metallic
and roughness bytes must change the shading, and a one-layer texture
must give the plain vertex-normal result.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures

W, H = 64, 64


def _quad_scene(handed: bool, nmr_texel: int | None):
    sc = S.build_patches(np.array([[0.0, 0.0, -2.0]]), np.eye(3)[None], np.array([3.0]), np.zeros(1, np.uint32),
                         [(0, True, 255)], [])
    albedo = np.full((16, 16), 0xFFB0B0B0, np.uint32)
    layers = [albedo] if nmr_texel is None else [albedo, np.full((16, 16), nmr_texel, np.uint32)]
    tex = textures.create(layers)
    hand = np.array([np.uint64(0xFFFFFFFFFFFFFFFF) if handed else np.uint64(0)], np.uint64)
    return dataclasses.replace(sc, textures=[tex], tangent_handedness=hand)


def _lit(sc, light_dir):
    view = S.make_view(W, H, count=1)
    ofb, _, b = oracle.render_frame(sc, view)
    assert (ofb.depth() > 0).all(), "the quad must cover the frame"
    lighting = S.default_lighting(view)
    d = np.asarray(light_dir, np.float64)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    lighting.lights = [S.Light(0, direction=tuple(d.tolist()), color=(1.0, 1.0, 1.0), intensity=1500.0)]
    rgba = oracle.resolve_lit(ofb, b, view.object_to_clip, lighting)
    return rgba.view(np.uint8).reshape(H, W, 4)[..., :3].astype(np.float64).mean()


# normal bytes (128, 218): tangent-space (0.004, 0.710, 0.704); metallic 0, roughness 255
LEAN_TO_B = 128 | (218 << 8) | (0 << 16) | (255 << 24)


@pytest.mark.parametrize("handed", [False, True])
def test_bitangent_orientation_and_handedness(handed):
    sc = _quad_scene(handed, LEAN_TO_B)
    from_plus_y = _lit(sc, (0.0, -1.0, -1.0))      # light travels towards -y-z: comes from +y+z
    from_minus_y = _lit(sc, (0.0, 1.0, -1.0))
    bright, dark = (from_minus_y, from_plus_y) if handed else (from_plus_y, from_minus_y)
    assert bright > 2.0 * dark + 20, (handed, from_plus_y, from_minus_y)


def test_one_layer_texture_uses_vertex_normal():
    plain = _quad_scene(False, None)
    a = _lit(plain, (0.0, -1.0, -1.0))
    b = _lit(plain, (0.0, 1.0, -1.0))
    assert abs(a - b) < 3.0        # the vertex normal (+z) sees both lights alike


def test_metallic_and_roughness_bytes():
    base = _lit(_quad_scene(False, LEAN_TO_B), (0.0, -1.0, -1.0))
    metal = _lit(_quad_scene(False, (LEAN_TO_B & 0xFF00FFFF) | (255 << 16)), (0.0, -1.0, -1.0))
    assert metal < base - 10       # a metal loses the Lambert term of its grey albedo
