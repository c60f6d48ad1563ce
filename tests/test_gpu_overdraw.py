"""Parity of the overdraw pipeline (OverdrawShader: FS_Overdraw,
Shading.cpp:333-342, 656) and its debug layers (ResolveDebug OverdrawPixel /
OverdrawQuad, Shading.cpp:759-766) with the CPU oracle: counter words and
depth bit-exact, debug colours bit-exact."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu


def _oracle(sc, view, clear=0, draws=((0, None),)):
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(view.width, view.height)
    fb.clear(clear, 0.0)
    total = {}
    for first, count in draws:
        count = sc.count - first if count is None else count
        sub, _ = oracle.cull(b, view.frustum_planes(), first, count)
        ctr, _ = oracle.draw(fb, b, view.object_to_clip, count, first, sub, pipeline=oracle.PIPELINE_OVERDRAW)
        for k, v in ctr.items():
            total[k] = total.get(k, 0) + v
    return fb, b, total


def _gpu(ctx, sc, view, clear=0, draws=((0, None),)):
    import glimpsw_amd as g

    ds = g.Scene(ctx, sc)
    fb = g.Framebuffer(ctx, view.width, view.height)
    fb.clear(clear, 0.0)
    ctx.reset_counters()
    for first, count in draws:
        count = sc.count - first if count is None else count
        fb.draw(ds, view.object_to_clip, count, first, view.frustum_planes(), g.GLIMP_DRAW_COUNTERS,
                pipeline=g.GLIMP_PIPELINE_OVERDRAW)
    return ds, fb


def _check(label, ctx, sc, view, clear=0, draws=((0, None),)):
    import glimpsw_amd as g

    print(f"[overdraw] {label}", flush=True)
    ofb, b, octr = _oracle(sc, view, clear, draws)
    ds, fb = _gpu(ctx, sc, view, clear, draws)
    try:
        words, depth = fb.read_layer(0), fb.read_layer(1)
        ow, od = ofb.pixels(0), ofb.pixels(1)
        bad = (words != ow) | (depth != od)
        if bad.any():
            ys, xs = np.nonzero(bad)
            rows = [f"({x},{y}) gpu {words[y, x]:#x} {depth[y, x]:#x} oracle {ow[y, x]:#x} {od[y, x]:#x}"
                    for y, x in list(zip(ys, xs))[:5]]
            pytest.fail(f"{label}: {bad.sum()} pixels differ\n" + "\n".join(rows))
        ctr = ctx.counters()
        assert ctr["fragments_covered"] == octr["fragments_covered"]
        assert ctr["fragments_passed"] == octr["fragments_passed"]
        assert ctr["dropped"] == 0
        for mode, layer in ((g.GLIMP_RESOLVE_OVERDRAW_PIXEL, oracle.DEBUG_OVERDRAW_PIXEL),
                            (g.GLIMP_RESOLVE_OVERDRAW_QUAD, oracle.DEBUG_OVERDRAW_QUAD)):
            fb.resolve(ds, view.object_to_clip, mode)
            got = fb.read_color()
            ofb2, b2, _ = _oracle(sc, view, clear, draws)
            want = oracle.resolve_debug(ofb2, b2, view.object_to_clip, layer)
            assert (got == want).all(), (label, mode, int((got != want).sum()))
        return words
    finally:
        fb.close()
        ds.close()


def test_overdraw_with_clipping(gpu_ctx):
    sc = T.small_scene(512, seed_offset=3, spread=6.0, near_frac=0.1)
    view = S.make_view(320, 240, count=sc.count)
    words = _check("clip scene", gpu_ctx, sc, view)
    assert (words >> 16).max() > 3


def test_overdraw_several_draws_and_clear_colour(gpu_ctx):
    sc = T.small_scene(384, seed_offset=5, spread=8.0)
    view = S.make_view(256, 192, count=sc.count)
    _check("three draws", gpu_ctx, sc, view, clear=0x00020003, draws=((0, 100), (100, 150), (250, None)))
    _check("saturating clear", gpu_ctx, sc, view, clear=0xFFFDFFFE)


def test_overdraw_1080p_frame(gpu_ctx):
    sc, view = S.checker_scene(2048, 1920, 1080)
    _check("1080p", gpu_ctx, sc, view)


def test_pipelines_do_not_mix(gpu_ctx):
    import glimpsw_amd as g

    sc = T.small_scene(64, seed_offset=1)
    view = S.make_view(128, 96, count=sc.count)
    ds, fb = _gpu(gpu_ctx, sc, view)
    try:
        with pytest.raises(g.GlimpError):
            fb.draw(ds, view.object_to_clip, sc.count, 0, view.frustum_planes())
        with pytest.raises(g.GlimpError):
            fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
        fb.clear(0, 0.0)
        fb.draw(ds, view.object_to_clip, sc.count, 0, view.frustum_planes())
        with pytest.raises(g.GlimpError):
            fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_OVERDRAW_PIXEL)
    finally:
        fb.close()
        ds.close()
