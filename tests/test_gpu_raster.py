"""Parity of the HIP vis-buffer path with the CPU oracle (bit-exact ids and
depth), on seeded scenes and the edge cases the reference's draw loop has:
ragged meshlets, guard-band clipping, back faces, several draws into one
framebuffer, tiny and maximum-size framebuffers, empty draws."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T

pytestmark = pytest.mark.gpu


def _gpu_frame(ctx, sc, view, cull=True, flags=0, draws=None, bin_size=0, bin_capacity=0):
    import glimpsw_amd as g

    ds = g.Scene(ctx, sc)
    fb = g.Framebuffer(ctx, view.width, view.height)
    fb.clear(0, 0.0)
    planes = view.frustum_planes() if cull else None
    for off, cnt in (draws or [(view.meshlet_offset, view.meshlet_count)]):
        fb.draw(ds, view.object_to_clip, cnt, off, planes if cull else None, flags, bin_size=bin_size,
                bin_capacity=bin_capacity)
    ids, depth = fb.read()
    fb.close()
    ds.close()
    return ids, depth


def _oracle_frame(sc, view, cull=True, clipping=True, guardband=True, draws=None):
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(view.width, view.height)
    fb.clear(0, 0.0)
    ctr = None
    for off, cnt in (draws or [(view.meshlet_offset, view.meshlet_count)]):
        bitmap = None
        if cull:
            bitmap, _ = oracle.cull(b, view.frustum_planes(), off, cnt)
        ctr, _ = oracle.draw(fb, b, view.object_to_clip, cnt, off, bitmap, clipping=clipping, guardband=guardband)
    return fb.ids(), fb.depth(), ctr


def _assert_same(g_ids, g_depth, o_ids, o_depth, label):
    print(f"[parity] {label}: covered={(o_depth > 0).sum()}", flush=True)
    bad = (g_ids != o_ids) | (g_depth.view(np.uint32) != o_depth.view(np.uint32))
    if bad.any():
        ys, xs = np.nonzero(bad)
        msg = [f"{label}: {bad.sum()} pixels differ, first at (x={xs[0]}, y={ys[0]})"]
        for y, x in list(zip(ys, xs))[:5]:
            msg.append(f"  ({x},{y}) gpu id={g_ids[y, x]} d={g_depth[y, x]!r} oracle id={o_ids[y, x]} d={o_depth[y, x]!r}")
        pytest.fail("\n".join(msg))


def test_checker_scene_frame(gpu_ctx):
    sc, view = S.checker_scene()
    g = _gpu_frame(gpu_ctx, sc, view)
    o = _oracle_frame(sc, view)
    assert (o[1] > 0).sum() > 10000
    _assert_same(*g, o[0], o[1], "checker_scene")


def test_checker_scene_no_cull(gpu_ctx):
    sc, view = S.checker_scene()
    g = _gpu_frame(gpu_ctx, sc, view, cull=False)
    o = _oracle_frame(sc, view, cull=False)
    _assert_same(*g, o[0], o[1], "checker_scene no cull")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_clipping_scene(gpu_ctx, seed):
    sc = T.small_scene(384, seed_offset=seed, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    assert o[2]["triangles_clipped"] > 0
    g = _gpu_frame(gpu_ctx, sc, view)
    _assert_same(*g, o[0], o[1], f"clip seed {seed}")


def test_clipping_disabled_and_no_guardband(gpu_ctx):
    import glimpsw_amd as g

    sc = T.small_scene(256, seed_offset=5, spread=5.0, near_frac=0.3)
    view = S.make_view(256, 192, count=sc.count)
    o = _oracle_frame(sc, view, clipping=False)
    gg = _gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_NO_CLIPPING)
    _assert_same(*gg, o[0], o[1], "no clipping")
    o = _oracle_frame(sc, view, guardband=False)
    gg = _gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_NO_GUARDBAND)
    _assert_same(*gg, o[0], o[1], "no guardband")


def test_ragged_meshlets(gpu_ctx):
    sc = T.ragged(T.small_scene(512, seed_offset=11, spread=8.0, near_frac=0.1))
    view = S.make_view(400, 300, count=sc.count)
    o = _oracle_frame(sc, view)
    g = _gpu_frame(gpu_ctx, sc, view)
    _assert_same(*g, o[0], o[1], "ragged")


def test_double_sided(gpu_ctx):
    sc = T.small_scene(300, seed_offset=13, spread=7.0, near_frac=0.1, double_sided=True)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    g = _gpu_frame(gpu_ctx, sc, view)
    _assert_same(*g, o[0], o[1], "double sided")


def test_multiple_draws_order(gpu_ctx):
    """Later draws come later in draw order: depth ties go to the earlier
    draw, ids carry each draw's meshlet offset."""
    sc = T.small_scene(400, seed_offset=17, spread=6.0)
    # duplicate the meshlets so the second draw is coplanar with the first
    import dataclasses
    dup = dataclasses.replace(
        sc, bounds=np.concatenate([sc.bounds] * 2), num_vertices=np.concatenate([sc.num_vertices] * 2),
        num_triangles=np.concatenate([sc.num_triangles] * 2), material_id=np.concatenate([sc.material_id] * 2),
        tangent_handedness=np.concatenate([sc.tangent_handedness] * 2),
        positions=np.concatenate([sc.positions] * 2), texcoords=np.concatenate([sc.texcoords] * 2),
        normal_tangents=np.concatenate([sc.normal_tangents] * 2), indices=np.concatenate([sc.indices] * 2))
    view = S.make_view(320, 240, count=dup.count)
    draws = [(400, 150), (0, 400), (150, 250)]
    o = _oracle_frame(dup, view, draws=draws)
    g = _gpu_frame(gpu_ctx, dup, view, draws=draws)
    _assert_same(*g, o[0], o[1], "multi draw")
    ids = g[0][g[1] > 0]
    assert (ids // 128 >= 400).any() and (ids // 128 < 400).any()


def test_tiny_and_max_framebuffer(gpu_ctx):
    sc, _ = S.checker_scene(256)
    for w, h in [(4, 4), (8, 12), (2896, 1628)]:
        view = S.make_view(w, h, count=sc.count)
        o = _oracle_frame(sc, view)
        g = _gpu_frame(gpu_ctx, sc, view)
        _assert_same(*g, o[0], o[1], f"{w}x{h}")


def test_watertight_grid(gpu_ctx):
    import glimpsw_amd as g

    sc = T.fullscreen_grid(256, 200)
    view = S.make_view(256, 200, count=sc.count)
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(256, 200)
    fb.clear(0, 0.0)
    _, cov = oracle.draw(fb, b, view.object_to_clip, sc.count, coverage=True)
    assert (cov == 1).all(), "oracle: grid must cover every pixel exactly once"
    gg = _gpu_frame(gpu_ctx, sc, view, cull=False)
    _assert_same(*gg, fb.ids(), fb.depth(), "grid")


def test_empty_draw_and_clear_values(gpu_ctx):
    import glimpsw_amd as g

    sc, view = S.checker_scene(64)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, 64, 48)
    fb.clear(0xFF000000, 0.0)
    fb.draw(ds, view.object_to_clip, 0)
    ids, depth = fb.read()
    assert (ids == 0xFF000000).all() and (depth == 0).all()
    with pytest.raises(g.GlimpError):
        fb.clear(0, -1.0)
    with pytest.raises(g.GlimpError):
        g.Framebuffer(gpu_ctx, 30, 32)
    with pytest.raises(g.GlimpError):
        fb.draw(ds, view.object_to_clip, 65)
    fb.close()
    ds.close()


@pytest.mark.parametrize("bin_capacity", [0, 4])
def test_counters_match_oracle(gpu_ctx, bin_capacity):
    """Capacity 4 overflows most bins: the (triangle, bin) pairs past a bin's
    capacity are drawn by the overflow raster instead of the bin raster, and
    each fragment is still counted once."""
    import glimpsw_amd as g

    sc = T.small_scene(384, seed_offset=2, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    gpu_ctx.reset_counters()
    _gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_COUNTERS, bin_size=16, bin_capacity=bin_capacity)
    c = gpu_ctx.counters()
    for k in ("triangles_processed", "triangles_rasterized", "triangles_clipped", "fragments_covered"):
        assert c[k] == o[2][k], (k, c[k], o[2][k])
    assert c["dropped"] == 0


@pytest.mark.parametrize("direct", [False, True])
def test_full_clip_queue_is_exact(gpu_ctx, direct):
    """600 identical small patches through the eye, sheared so the near
    plane cuts a diagonal of each (7,200 clipped triangles in the oracle),
    queue more clipped triangles than the clip queue holds (4096 or two per
    meshlet) and produce more fan triangles than the record region holds:
    the triangles without queue room are clipped by their setup wave, the fan
    triangles without a record slot are drawn at once, and the frame equals
    the oracle's bit for bit with nothing dropped, on the binned and the
    direct route. The next, ordinary draw (whose scratch counters the full
    one must leave cleared) is exact too."""
    import glimpsw_amd as g

    n = 600
    R = np.tile(np.array([[0.7, 0.0, 0.0], [0.0, 0.7, 0.0], [0.7, 0.7, 1.0]]), (n, 1, 1))
    sc = S.build_patches(np.tile([[0.0, 0.0, -S.NEAR_Z]], (n, 1)), R, np.full(n, 0.05), np.zeros(n, np.uint32),
                         [(0, True, 255)], [textures.checker_texture(64)])
    view = S.make_view(320, 240, count=n)
    o = _oracle_frame(sc, view, cull=False)
    assert o[2]["triangles_clipped"] > 4096
    gpu_ctx.reset_counters()
    flags = g.GLIMP_DRAW_COUNTERS | (g.GLIMP_DRAW_DIRECT if direct else 0)
    got = _gpu_frame(gpu_ctx, sc, view, cull=False, flags=flags)
    c = gpu_ctx.counters()
    print(f"[raster] full clip queue: clipped {c['triangles_clipped']}, dropped {c['dropped']}", flush=True)
    _assert_same(*got, o[0], o[1], "full clip queue")
    assert c["dropped"] == 0
    for k in ("triangles_processed", "triangles_rasterized", "triangles_clipped", "fragments_covered"):
        assert c[k] == o[2][k], (k, c[k], o[2][k])
    sc2, view2 = S.checker_scene()
    _assert_same(*_gpu_frame(gpu_ctx, sc2, view2), *_oracle_frame(sc2, view2)[:2], "after a full clip queue")


@pytest.mark.slow
def test_frame_1080p(gpu_ctx):
    sc, view = S.frame_1080p()
    o = _oracle_frame(sc, view)
    g = _gpu_frame(gpu_ctx, sc, view)
    _assert_same(*g, o[0], o[1], "frame_1080p")


@pytest.mark.slow
def test_frame_1080p_crossers_inside(gpu_ctx):
    """The 1080p frame with its near-plane crossers reaching into the frustum
    (hundreds of clipped triangles covering much of the screen), binned at
    32-pixel bins and on the direct route, bit-exact with the oracle."""
    import glimpsw_amd as g

    sc, view = S.frame_1080p(crossers_inside=True)
    o = _oracle_frame(sc, view)
    assert o[2]["triangles_clipped"] > 300
    _assert_same(*_gpu_frame(gpu_ctx, sc, view, bin_size=32), o[0], o[1], "crossers inside, binned")
    _assert_same(*_gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_DIRECT), o[0], o[1], "crossers inside, direct")


@pytest.mark.parametrize("bin_size", [8, 16, 32, 64])
def test_bin_sizes(gpu_ctx, bin_size):
    sc = T.small_scene(384, seed_offset=1, spread=6.0, near_frac=0.25)
    view = S.make_view(324, 244, count=sc.count)   # partial bins at the right/bottom edges
    o = _oracle_frame(sc, view)
    g = _gpu_frame(gpu_ctx, sc, view, bin_size=bin_size)
    _assert_same(*g, o[0], o[1], f"bin {bin_size}")


@pytest.mark.parametrize("cap", [1, 3, 16])
def test_bin_spill_bit_identical(gpu_ctx, cap):
    """Bins far too small: most (triangle, bin) pairs overflow to the
    overflow raster, whose spill keys the bins merge, and the vis-buffer must
    not change."""
    sc, view = S.checker_scene()
    o = _oracle_frame(sc, view)
    g = _gpu_frame(gpu_ctx, sc, view, bin_size=16, bin_capacity=cap)
    _assert_same(*g, o[0], o[1], f"capacity {cap}")


def test_overflow_leaves_spill_keys_cleared(gpu_ctx):
    """An overflowing draw must leave its framebuffer's spill keys zero: a
    second scene drawn into the same framebuffer right after (its bins also
    overflow, so they merge the spill keys) must equal the oracle's frame of
    that scene alone, at every bin size."""
    import glimpsw_amd as g

    sc1, view = S.checker_scene()
    sc2 = T.small_scene(300, seed_offset=5, spread=6.0, near_frac=0.2)
    view2 = S.make_view(view.width, view.height, count=sc2.count)
    o = _oracle_frame(sc2, view2)
    d1, d2 = g.Scene(gpu_ctx, sc1), g.Scene(gpu_ctx, sc2)
    fb = g.Framebuffer(gpu_ctx, view.width, view.height)
    for bs in (8, 16, 32, 64):
        fb.clear(0, 0.0)
        fb.draw(d1, view.object_to_clip, sc1.count, 0, view.frustum_planes(), 0, bin_size=bs, bin_capacity=2)
        fb.read()
        fb.clear(0, 0.0)
        fb.draw(d2, view2.object_to_clip, sc2.count, 0, view2.frustum_planes(), 0, bin_size=bs, bin_capacity=2)
        ids, depth = fb.read()
        _assert_same(ids, depth, o[0], o[1], f"second overflowing draw, bin {bs}")
    fb.close()
    d1.close()
    d2.close()


def test_direct_path(gpu_ctx):
    import glimpsw_amd as g

    sc = T.small_scene(384, seed_offset=0, spread=6.0, near_frac=0.25)
    view = S.make_view(320, 240, count=sc.count)
    o = _oracle_frame(sc, view)
    gg = _gpu_frame(gpu_ctx, sc, view, flags=g.GLIMP_DRAW_DIRECT)
    _assert_same(*gg, o[0], o[1], "direct")


def test_mixed_direct_and_binned_draws(gpu_ctx):
    """A binned draw after a direct one starts its bins from the keys
    already in the vis-buffer."""
    import glimpsw_amd as g

    sc = T.small_scene(400, seed_offset=21, spread=6.0)
    view = S.make_view(320, 240, count=sc.count)
    draws = [(0, 200), (200, 200)]
    o = _oracle_frame(sc, view, draws=draws)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, 320, 240)
    fb.clear(0, 0.0)
    fb.draw(ds, view.object_to_clip, 200, 0, view.frustum_planes(), g.GLIMP_DRAW_DIRECT)
    fb.draw(ds, view.object_to_clip, 200, 200, view.frustum_planes(), 0)
    ids, depth = fb.read()
    _assert_same(ids, depth, o[0], o[1], "mixed")
    fb.close()
    ds.close()


def test_growing_bin_grid_starts_cleared():
    """Draws on one context with ever more bins: each draw regrows the bin
    count buffer, which must start at zero even when the allocator hands back
    memory just freed with other contents (here a framebuffer full of keys,
    freed right before)."""
    import glimpsw_amd as g

    ctx = g.Context(0)
    try:
        sc = T.small_scene(256, seed_offset=7, spread=6.0, near_frac=0.1)
        ds = g.Scene(ctx, sc)
        for w, h in ((64, 48), (128, 96), (200, 180), (320, 240), (480, 368)):
            junk = g.Framebuffer(ctx, w * 2, h * 2)
            junk.clear(0, 0.0)
            jv = S.make_view(w * 2, h * 2, count=sc.count)
            junk.draw(ds, jv.object_to_clip, sc.count, 0, jv.frustum_planes())
            junk.read()
            junk.close()
            view = S.make_view(w, h, count=sc.count)
            o = _oracle_frame(sc, view)
            fb = g.Framebuffer(ctx, w, h)
            fb.clear(0, 0.0)
            fb.draw(ds, view.object_to_clip, sc.count, 0, view.frustum_planes(), 0, bin_size=8, bin_capacity=64)
            ids, depth = fb.read()
            fb.close()
            _assert_same(ids, depth, o[0], o[1], f"{w}x{h}")
        ds.close()
    finally:
        ctx.close()
