"""Parity of the HIP resolve with the CPU oracle: albedo (ResolveDebug
BaseColor) bit-exact, lit resolve within 1/255 per channel (the oracle and
the kernel evaluate the same operation order, so it is expected exact)."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu


def _frames(ctx, sc, view, lit):
    import glimpsw_amd as g

    ds = g.Scene(ctx, sc)
    fb = g.render_frame(ctx, ds, view)
    fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
    g_alb = fb.read_color()
    lighting = S.default_lighting(view)
    fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)
    g_lit = fb.read_color()
    fb.close()
    ds.close()

    ofb, _, b = oracle.render_frame(sc, view)
    o_alb = oracle.resolve_albedo(ofb, b, view.object_to_clip)
    ofb2, _, _ = oracle.render_frame(sc, view, b)
    o_lit = oracle.resolve_lit(ofb2, b, view.object_to_clip, lighting)
    return g_alb, o_alb, g_lit, o_lit


def _check(g_alb, o_alb, g_lit, o_lit, label):
    print(f"[resolve] {label}", flush=True)
    bad = g_alb != o_alb
    assert not bad.any(), f"{label}: albedo differs at {bad.sum()} pixels, first {np.argwhere(bad)[:3].tolist()}"
    gl = g_lit.view(np.uint8).reshape(*g_lit.shape, 4).astype(int)
    ol = o_lit.view(np.uint8).reshape(*o_lit.shape, 4).astype(int)
    diff = np.abs(gl - ol).max()
    assert diff <= 1, f"{label}: lit differs by {diff}/255"


def test_checker_scene(gpu_ctx):
    sc, view = S.checker_scene()
    _check(*_frames(gpu_ctx, sc, view, True), "checker_scene")


def test_clip_scene(gpu_ctx):
    sc = T.small_scene(384, seed_offset=0, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    _check(*_frames(gpu_ctx, sc, view, True), "clip scene")


def test_frame_1080p(gpu_ctx):
    sc, view = S.frame_1080p()
    _check(*_frames(gpu_ctx, sc, view, True), "frame_1080p")


def test_resolve_after_multiple_draws(gpu_ctx):
    """Surface ids of later draws carry their meshlet offsets."""
    import glimpsw_amd as g

    sc = T.small_scene(300, seed_offset=4, spread=6.0)
    view = S.make_view(320, 240, count=sc.count)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, 320, 240)
    fb.clear(0, 0.0)
    fb.draw(ds, view.object_to_clip, 120, 180)
    fb.draw(ds, view.object_to_clip, 180, 0)
    fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
    got = fb.read_color()
    fb.close()
    ds.close()
    b = oracle.SceneBinding(sc)
    ofb = oracle.OracleFramebuffer(320, 240)
    ofb.clear(0, 0.0)
    oracle.draw(ofb, b, view.object_to_clip, 120, 180)
    oracle.draw(ofb, b, view.object_to_clip, 180, 0)
    want = oracle.resolve_albedo(ofb, b, view.object_to_clip)
    assert (got == want).all()


def test_resolve_without_draw_is_background(gpu_ctx):
    import glimpsw_amd as g

    sc, view = S.checker_scene(16)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, 16, 8)
    fb.clear(0, 0.0)
    fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
    got = fb.read_color()
    ys, xs = np.mgrid[0:8, 0:16]
    want = np.where((((xs & ~3) ^ (ys & ~3)) & 4) != 0, 0xFFA0A0A0, 0xFFFFFFFF).astype(np.uint32)
    assert (got == want).all()
    fb.close()
    ds.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_normal_mapped_scene(gpu_ctx, mixed):
    """Textures with a normal/metallic/roughness layer (ResolveSurface,
    Shading.cpp:541-577), alone and alternating with one-layer textures in
    the same tiles; seeded tangent handedness bits."""
    sc = T.normal_mapped_scene(384, seed_offset=3 + mixed, near_frac=0.2, mixed=mixed)
    view = S.make_view(320, 240, count=sc.count)
    _check(*_frames(gpu_ctx, sc, view, True), f"normal-mapped mixed={mixed}")


def test_light_sprites(gpu_ctx):
    """Point and spot light sprites over the lit image (Shading.cpp:689-727):
    a light in front of the surfaces near it and one partly hidden behind a
    patch, as a point and as a spot light."""
    import glimpsw_amd as g
    from tests.test_oracle_sprites import _light_positions

    sc, view = S.checker_scene(width=1280, height=960)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.render_frame(gpu_ctx, ds, view)
    ofb, _, b = oracle.render_frame(sc, view)
    positions = list(_light_positions(sc, view, ofb.depth()))
    lighting = S.default_lighting(view)
    for pos in positions:
        for kind in (1, 2):
            lighting.lights[1] = S.Light(kind, position=pos, direction=(0.0, 0.0, -1.0), color=(1.2, 0.7, 0.3),
                                         intensity=4000.0, radius=30.0)
            fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)
            got = fb.read_color()                       # the vis-buffer stays; resolve writes colour only
            ofb2, _, _ = oracle.render_frame(sc, view, b)
            want = oracle.resolve_lit(ofb2, b, view.object_to_clip, lighting)
            gl = got.view(np.uint8).reshape(*got.shape, 4).astype(int)
            ol = want.view(np.uint8).reshape(*want.shape, 4).astype(int)
            print(f"[resolve] sprite at {pos} type {kind}: max diff {np.abs(gl - ol).max()}", flush=True)
            assert np.abs(gl - ol).max() <= 1
    fb.close()
    ds.close()


@pytest.mark.parametrize("layer", ["normals", "metallic_roughness", "meshlet_id", "triangle_id"])
def test_debug_layers(gpu_ctx, layer):
    """The other ResolveDebug layers (Shading.cpp:734-773) on a normal-mapped
    scene mixing one- and two-layer textures, bit-exact with the oracle."""
    import glimpsw_amd as g

    modes = {"normals": (g.GLIMP_RESOLVE_NORMALS, oracle.DEBUG_NORMALS),
             "metallic_roughness": (g.GLIMP_RESOLVE_METALLIC_ROUGHNESS, oracle.DEBUG_METALLIC_ROUGHNESS),
             "meshlet_id": (g.GLIMP_RESOLVE_MESHLET_ID, oracle.DEBUG_MESHLET_ID),
             "triangle_id": (g.GLIMP_RESOLVE_TRIANGLE_ID, oracle.DEBUG_TRIANGLE_ID)}
    gmode, omode = modes[layer]
    sc = T.normal_mapped_scene(384, seed_offset=6, near_frac=0.2, mixed=True)
    view = S.make_view(320, 240, count=sc.count)
    lighting = S.default_lighting(view)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.render_frame(gpu_ctx, ds, view)
    fb.resolve(ds, view.object_to_clip, gmode, lighting)
    got = fb.read_color()
    fb.close()
    ds.close()
    ofb, _, b = oracle.render_frame(sc, view)
    want = oracle.resolve_debug(ofb, b, view.object_to_clip, omode, lighting.object_to_world)
    bad = got != want
    print(f"[resolve] debug layer {layer}: {bad.sum()} pixels differ", flush=True)
    assert not bad.any(), np.argwhere(bad)[:3].tolist()


@pytest.mark.parametrize("size", [(320, 240, 128), (1920, 1080, 256)])
def test_skybox(gpu_ctx, size):
    """The lit resolve with an R11G11B10f octahedral skybox for the pixels
    nothing covers (Shading.cpp:659-678): sky pixels bit-exact, covered
    pixels within 1/255, with the light sprites on."""
    import glimpsw_amd as g
    from glimpsw_amd import textures

    W, H, tsize = size
    sc, view = S.checker_scene(1024 if W > 1000 else 192, W, H)
    sky = textures.sky_texture(tsize)
    ds = g.Scene(gpu_ctx, sc)
    dsky = g.Texture(gpu_ctx, sky)
    fb = g.render_frame(gpu_ctx, ds, view)
    lighting = S.default_lighting(view)
    fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting, skybox=dsky)
    got = fb.read_color()
    ids, depth = fb.read()
    fb.close()
    dsky.close()
    ds.close()
    lighting.skybox = sky
    ofb, _, b = oracle.render_frame(sc, view)
    want = oracle.resolve_lit(ofb, b, view.object_to_clip, lighting)
    skyp = depth <= 0
    print(f"[resolve] skybox {W}x{H}: {skyp.sum()} sky pixels", flush=True)
    assert skyp.sum() > 1000
    bad = (got != want) & skyp
    assert not bad.any(), f"{bad.sum()} sky pixels differ, first {np.argwhere(bad)[:3].tolist()}"
    gl = got.view(np.uint8).reshape(*got.shape, 4).astype(int)
    ol = want.view(np.uint8).reshape(*want.shape, 4).astype(int)
    assert np.abs(gl - ol).max() <= 1
