"""Parity of the deferred pipeline (DeferredShader: FS_EncodeGBuffer,
Shading.cpp:344-407, 655) with the CPU oracle: base colour, depth and the
packed normal / metallic / roughness layer bit-exact, with normal-mapped,
mixed and alpha-tested materials, guard-band clipping, several draws and a
second frame over an uncleared layer 2."""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu

ROT = np.array([[np.cos(0.5), 0.0, -np.sin(0.5)], [0.0, 1.0, 0.0], [np.sin(0.5), 0.0, np.cos(0.5)]],
               np.float32).reshape(9)          # mat3, column-major


def _oracle_frames(sc, view, frames, o2w):
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(view.width, view.height, layers=3)
    out = []
    for clear, draws in frames:
        fb.clear(clear, 0.0)
        for first, count in draws:
            bm, _ = oracle.cull(b, view.frustum_planes(), first, count)
            oracle.draw(fb, b, view.object_to_clip, count, first, bm, pipeline=oracle.PIPELINE_DEFERRED,
                        object_to_world=o2w)
        out.append([fb.pixels(k).copy() for k in range(3)])
    return out


def _check(label, ctx, sc, view, frames=None, o2w=None):
    import glimpsw_amd as g

    frames = frames or [(0, [(0, sc.count)])]
    o2w = np.eye(3, dtype=np.float32).reshape(9) if o2w is None else o2w
    print(f"[deferred] {label}", flush=True)
    want = _oracle_frames(sc, view, frames, o2w)
    ds = g.Scene(ctx, sc)
    fb = g.Framebuffer(ctx, view.width, view.height)
    try:
        for f, (clear, draws) in enumerate(frames):
            fb.clear(clear, 0.0)
            for first, count in draws:
                fb.draw(ds, view.object_to_clip, count, first, view.frustum_planes(), pipeline=g.GLIMP_PIPELINE_DEFERRED,
                        object_to_world=o2w)
            for k in range(3):
                got = fb.read_layer(k)
                bad = got != want[f][k]
                if bad.any():
                    ys, xs = np.nonzero(bad)
                    rows = [f"({x},{y}) gpu {got[y, x]:#010x} oracle {want[f][k][y, x]:#010x}"
                            for y, x in list(zip(ys, xs))[:5]]
                    pytest.fail(f"{label} frame {f} layer {k}: {bad.sum()} pixels differ\n" + "\n".join(rows))
        assert ctx.counters()["dropped"] == 0
        return want
    finally:
        fb.close()
        ds.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_normal_mapped_with_clipping(gpu_ctx, mixed):
    sc = T.normal_mapped_scene(384, seed_offset=4, near_frac=0.15, mixed=mixed)
    view = S.make_view(320, 240, count=sc.count)
    want = _check(f"normal mapped mixed={mixed}", gpu_ctx, sc, view, o2w=ROT)
    assert (want[0][2] != 0).sum() > 5000


def test_alpha_tested_materials(gpu_ctx):
    sc = T.alpha_scene(384, seed_offset=6, spread=6.0, near_frac=0.1)
    view = S.make_view(256, 192, count=sc.count)
    _check("alpha scene", gpu_ctx, sc, view)


def test_several_draws_and_uncleared_layer2(gpu_ctx):
    sc = T.normal_mapped_scene(300, seed_offset=8, mixed=True)
    view = S.make_view(256, 192, count=sc.count)
    frames = [(0x10203040, [(0, 120), (120, 180)]), (0x05060708, [(150, 150)])]
    want = _check("two frames", gpu_ctx, sc, view, frames=frames, o2w=ROT)
    # the second frame draws fewer meshlets: layer 2 keeps pixels of the first
    kept = (want[1][1] == 0) & (want[1][2] != 0)
    assert kept.sum() > 100


def test_1080p_frame(gpu_ctx):
    sc, view = S.checker_scene(2048, 1920, 1080)
    _check("1080p checker", gpu_ctx, sc, view)


def test_untextured_ranges_are_refused(gpu_ctx):
    import glimpsw_amd as g

    sc = T.small_scene(32, seed_offset=2)
    mid = sc.material_id.copy()
    mid[5] = 0xFFFFFFFF
    sc = dataclasses.replace(sc, material_id=mid)
    view = S.make_view(128, 96, count=sc.count)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, view.width, view.height)
    try:
        fb.clear(0, 0.0)
        with pytest.raises(g.GlimpError):
            fb.draw(ds, view.object_to_clip, sc.count, 0, pipeline=g.GLIMP_PIPELINE_DEFERRED)
        fb.draw(ds, view.object_to_clip, 5, 0, pipeline=g.GLIMP_PIPELINE_DEFERRED)     # meshlets 0-4 are fine
        with pytest.raises(g.GlimpError):
            fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
        with pytest.raises(g.GlimpError):
            fb.draw(ds, view.object_to_clip, 5, 0)            # a vis-buffer draw over G-buffer layers
    finally:
        fb.close()
        ds.close()
