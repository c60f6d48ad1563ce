"""Parity of the alpha-tested vis-buffer path (FS_EncodeSurfaceId<true>,
Shading.cpp:309-331) with the CPU oracle: bit-exact ids and depth with
cutout materials on every raster route (binned, bin spill, direct, batch,
with guard-band clipping), and the resolve of such a frame."""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T

pytestmark = pytest.mark.gpu


def _gpu_frame(ctx, sc, view, flags=0, bin_size=0, bin_capacity=0):
    import glimpsw_amd as g

    ds = g.Scene(ctx, sc)
    fb = g.Framebuffer(ctx, view.width, view.height)
    fb.clear(0, 0.0)
    fb.draw(ds, view.object_to_clip, sc.count, 0, view.frustum_planes(), flags, bin_size=bin_size,
            bin_capacity=bin_capacity)
    ids, depth = fb.read()
    fb.close()
    ds.close()
    return ids, depth


def _oracle_frame(sc, view):
    ofb, _, _ = oracle.render_frame(sc, view)
    return ofb.ids(), ofb.depth()


def _assert_same(g, o, label):
    print(f"[alpha] {label}: covered={(o[1] > 0).sum()}", flush=True)
    bad = (g[0] != o[0]) | (g[1].view(np.uint32) != o[1].view(np.uint32))
    if bad.any():
        ys, xs = np.nonzero(bad)
        rows = [f"({x},{y}) gpu {g[0][y, x]} {g[1][y, x]!r} oracle {o[0][y, x]} {o[1][y, x]!r}"
                for y, x in list(zip(ys, xs))[:5]]
        pytest.fail(f"{label}: {bad.sum()} pixels differ\n" + "\n".join(rows))


def _plain(sc):
    return dataclasses.replace(sc, materials=[(t, ds, 255) for (t, ds, _) in sc.materials])


@pytest.mark.parametrize("seed", [0, 1])
def test_alpha_scene_with_clipping(gpu_ctx, seed):
    sc = T.alpha_scene(384, seed_offset=seed, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    # the cutout changes the frame: the test would pass vacuously otherwise
    assert (o[0] != _oracle_frame(_plain(sc), view)[0]).sum() > 500
    _assert_same(_gpu_frame(gpu_ctx, sc, view), o, f"binned seed {seed}")


def test_alpha_direct_and_spill(gpu_ctx):
    import glimpsw_amd as g

    sc = T.alpha_scene(384, seed_offset=2, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    _assert_same(_gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_DIRECT), o, "direct")
    _assert_same(_gpu_frame(gpu_ctx, sc, view, bin_size=16, bin_capacity=4), o, "spill (capacity 4)")


@pytest.mark.parametrize("bin_size", [8, 64])
def test_alpha_bin_sizes(gpu_ctx, bin_size):
    sc = T.alpha_scene(512, seed_offset=3)
    view = S.make_view(640, 480, count=sc.count)
    _assert_same(_gpu_frame(gpu_ctx, sc, view, bin_size=bin_size), _oracle_frame(sc, view), f"bin {bin_size}")


def test_alpha_cutoff_extremes(gpu_ctx):
    """Cutoff 0 passes every texel; cutoff 254 keeps only alpha-255 texels."""
    for cut in ((0, 0), (254, 254)):
        sc = T.alpha_scene(256, seed_offset=4, spread=6.0, cutoffs=cut)
        view = S.make_view(320, 240, count=sc.count)
        _assert_same(_gpu_frame(gpu_ctx, sc, view), _oracle_frame(sc, view), f"cutoffs {cut}")


def test_plain_then_alpha_draws(gpu_ctx):
    """Draws into one framebuffer: an opaque scene, then an alpha-tested one
    (the alpha side buffers appear between the two draws)."""
    import glimpsw_amd as g

    plain = T.small_scene(256, seed_offset=5, spread=6.0)
    cut = T.alpha_scene(256, seed_offset=6, spread=6.0, near_frac=0.2)
    view = S.make_view(320, 240)
    dp, dc = g.Scene(gpu_ctx, plain), g.Scene(gpu_ctx, cut)
    fb = g.Framebuffer(gpu_ctx, 320, 240)
    fb.clear(0, 0.0)
    fb.draw(dp, view.object_to_clip, plain.count, 0, view.frustum_planes())
    fb.draw(dc, view.object_to_clip, cut.count, 0, view.frustum_planes())
    got = fb.read()
    fb.close()
    dp.close()
    dc.close()
    # oracle: both scenes into one framebuffer, draw-order ranks continue
    ofb = oracle.OracleFramebuffer(320, 240)
    ofb.clear(0, 0.0)
    for sc in (plain, cut):
        b = oracle.SceneBinding(sc)
        bm, _ = oracle.cull(b, view.frustum_planes(), 0, sc.count)
        oracle.draw(ofb, b, view.object_to_clip, sc.count, 0, bm)
    _assert_same(got, (ofb.ids(), ofb.depth()), "opaque then cutout draw")


@pytest.mark.parametrize("shape", [None, {"bin_size": 32, "bin_capacity": 512, "raster_threads": 1024},
                                   {"bin_size": 64, "bin_capacity": 4096, "raster_threads": 512}])
def test_alpha_batch(gpu_ctx, shape):
    """A batch mixing alpha-tested and opaque scenes equals its per-view
    draws and the oracle (default shape and the wide bin rasters)."""
    import glimpsw_amd as g

    scenes = [T.alpha_scene(300, seed_offset=7, spread=6.0, near_frac=0.2), T.small_scene(300, seed_offset=8)]
    dsc = [g.Scene(gpu_ctx, s) for s in scenes]
    views = []
    for i, (w, h) in enumerate([(320, 240), (256, 128), (128, 256), (64, 64)]):
        yaw = 0.3 * i
        v = S.make_view(w, h, view=S.view_matrix((0.0, 0.0, 0.0), yaw, 0.0), count=scenes[i % 2].count)
        views.append((i % 2, v))
    fbs = [g.Framebuffer(gpu_ctx, v.width, v.height) for _, v in views]
    items = [(dsc[si], fb, v) for (si, v), fb in zip(views, fbs)]
    b = g.Batch(gpu_ctx, items, shape=shape)
    b.launch_per_view()
    want = [fb.read() for fb in fbs]
    b.launch(graph=True)
    got = [fb.read() for fb in fbs]
    b.close()
    for i, ((si, v), w, gt) in enumerate(zip(views, want, got)):
        _assert_same(gt, w, f"batch view {i} vs per-view")
        ofb, _, _ = oracle.render_frame(scenes[si], v)
        _assert_same(gt, (ofb.ids(), ofb.depth()), f"batch view {i} vs oracle")
    for fb in fbs:
        fb.close()
    for d in dsc:
        d.close()


def test_alpha_resolve(gpu_ctx):
    """ResolveDebug(BaseColor) bit-exact and the lit resolve within 1/255 on
    a frame with cutout holes."""
    import glimpsw_amd as g

    sc = T.alpha_scene(384, seed_offset=9, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.render_frame(gpu_ctx, ds, view)
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
    assert (g_alb == o_alb).all()
    gl = g_lit.view(np.uint8).reshape(*g_lit.shape, 4).astype(int)
    ol = o_lit.view(np.uint8).reshape(*o_lit.shape, 4).astype(int)
    assert np.abs(gl - ol).max() <= 1


def test_alpha_frame_1080p(gpu_ctx):
    """The 1080p frame with half its materials turned into cutouts."""
    sc, view = S.frame_1080p()
    mats = [(i, False, 128 if i % 2 else 255) for i in range(len(sc.materials))]
    texs = [textures.cutout_texture(256, S.SEED_BASE + 800 + i) if i % 2 else t for i, t in enumerate(sc.textures)]
    sc = dataclasses.replace(sc, materials=mats, textures=texs)
    _assert_same(_gpu_frame(gpu_ctx, sc, view), _oracle_frame(sc, view), "1080p cutout")
