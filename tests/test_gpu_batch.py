"""Batched rendering: one launch per stage (and one hipGraph) over many views
of several scenes gives every view exactly the vis-buffer of a clear plus one
DrawMeshlets of that view, for every launch shape the autotuner may pick; a
small batch is also checked against the CPU oracle."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S

pytestmark = pytest.mark.gpu


def _small_batch_views(num_views=10):
    return S.batch_views(num_scenes=4, num_views=num_views, min_meshlets=100, max_meshlets=1500,
                     min_res=(64, 64), max_res=(480, 270))


class _Setup:
    def __init__(self, ctx, scenes, views):
        import glimpsw_amd as g

        self.g = g
        self.dscenes = [g.Scene(ctx, sc) for sc in scenes]
        self.fbs = [g.Framebuffer(ctx, v.width, v.height) for _, v in views]
        self.items = [(self.dscenes[si], fb, v) for (si, v), fb in zip(views, self.fbs)]

    def read_all(self):
        return [fb.read() for fb in self.fbs]

    def close(self):
        for fb in self.fbs:
            fb.close()
        for d in self.dscenes:
            d.close()


def _same(a, b, label):
    for i, ((ia, da), (ib, db)) in enumerate(zip(a, b)):
        bad = (ia != ib) | (da.view(np.uint32) != db.view(np.uint32))
        assert not bad.any(), f"{label}: view {i} differs at {int(bad.sum())} pixels"


@pytest.fixture(scope="module")
def small(gpu_ctx):
    scenes, views = _small_batch_views()
    st = _Setup(gpu_ctx, scenes, views)
    yield scenes, views, st
    st.close()


def test_batch_matches_per_view(gpu_ctx, small):
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    print(f"[batch] per-view vs batched, {len(views)} views", flush=True)
    b.launch_per_view()
    want = st.read_all()
    assert any((ids != 0).any() for ids, _ in want), "nothing drawn"
    b.launch(graph=False)
    _same(want, st.read_all(), "eager batch")
    b.launch(graph=True)
    _same(want, st.read_all(), "graph batch")
    b.close()


def test_batch_vs_oracle(gpu_ctx, small):
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    b.launch(graph=True)
    got = st.read_all()
    b.close()
    bindings = {}
    for i, (si, v) in enumerate(views[:6]):
        print(f"[batch] oracle view {i} {v.width}x{v.height} scene {si}", flush=True)
        if si not in bindings:
            bindings[si] = oracle.SceneBinding(scenes[si])
        ofb, _, _ = oracle.render_frame(scenes[si], v, bindings[si])
        ids, depth = ofb.ids(), ofb.depth()
        _same([(ids, depth)], [got[i]], f"oracle view {i}")


def test_graph_replays_with_work_between(gpu_ctx, small):
    """Every replay of the captured graph (with eager draws clobbering the
    framebuffers between replays) reproduces the per-view result."""
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    b.launch_per_view()
    want = st.read_all()
    for r in range(3):
        b.launch(graph=True)
        _same(want, st.read_all(), f"replay {r}")
        for (ds, fb, v) in st.items[:3]:
            fb.clear(0x123, 0.0)
            fb.draw(ds, v.object_to_clip, v.meshlet_count // 2, 0)
    b.close()


@pytest.mark.parametrize("bin_size,capacity,threads", [
    (8, 256, 128), (16, 64, 256), (32, 4096, 256), (64, 2048, 128), (64, 3, 256),
    (32, 4096, 512), (32, 1024, 1024), (64, 2048, 1024), (32, 3, 1024),
])
def test_batch_shapes(gpu_ctx, small, bin_size, capacity, threads):
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items, shape={"bin_size": bin_size, "bin_capacity": capacity,
                                         "raster_threads": threads})
    print(f"[batch] shape {bin_size}/{capacity}/{threads}", flush=True)
    b.launch_per_view()
    want = st.read_all()
    b.launch(graph=True)
    _same(want, st.read_all(), f"shape {bin_size}/{capacity}/{threads}")
    b.close()


def test_autotune_keeps_results(gpu_ctx, small):
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    b.launch_per_view()
    want = st.read_all()
    shape, ms = b.autotune()
    print(f"[batch] autotune -> {shape} {ms:.3f} ms", flush=True)
    assert shape["bin_size"] in (8, 16, 32, 64)
    assert 256 <= shape["bin_capacity"] <= 4096
    assert shape["raster_threads"] in (128, 256, 512, 1024)
    assert ms > 0.0
    gpu_ctx.reset_counters()
    b.launch(graph=True)
    _same(want, st.read_all(), "autotuned")
    assert gpu_ctx.counters()["dropped"] == 0
    b.close()


def test_batch_then_resolve(gpu_ctx, small):
    """A batched view resolves like the same view drawn on its own."""
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    b.launch(graph=True)
    ds, fb, v = st.items[1]
    fb.resolve(ds, v.object_to_clip, g.GLIMP_RESOLVE_ALBEDO)
    got = fb.read_color()
    b.close()
    ofb, _, bind = oracle.render_frame(scenes[views[1][0]], v)
    want = oracle.resolve_albedo(ofb, bind, v.object_to_clip)
    assert (got == want).all()


def test_empty_views_are_cleared(gpu_ctx):
    import glimpsw_amd as g

    scenes, views = S.batch_views(num_scenes=2, num_views=3, min_meshlets=50, max_meshlets=200,
                              min_res=(40, 24), max_res=(100, 60))
    st = _Setup(gpu_ctx, scenes, views)
    for _, fb, _ in st.items:
        fb.clear(0, 0.0)
        fb.draw(st.items[0][0], views[0][1].object_to_clip, views[0][1].meshlet_count, 0)
    items = [(ds, fb, S.View(v.width, v.height, v.proj, v.view, v.object_to_clip, 0, 0))
             for ds, fb, v in st.items]
    b = g.Batch(gpu_ctx, items, clear_depth=0.25)
    b.launch(graph=True)
    for ids, depth in st.read_all():
        assert (ids == 0).all() and (depth == np.float32(0.25)).all()
    b.close()
    st.close()


def test_batch_rejects_shared_framebuffer(gpu_ctx, small):
    scenes, views, st = small
    g = st.g
    items = [st.items[0], (st.items[1][0], st.items[0][1], st.items[1][2])]
    with pytest.raises(g.GlimpError):
        g.Batch(gpu_ctx, items)
    with pytest.raises(g.GlimpError):
        g.Batch(gpu_ctx, st.items[:2], shape={"bin_size": 12, "bin_capacity": 64})


def test_batch_with_resolve(gpu_ctx, small):
    """Resolves added to the batch run inside its launch and graph: every
    view's colour equals draw + resolve view by view, for lit (with light
    sprites) and albedo resolves, on every replay, with eager draws and
    resolves of the same framebuffers between the replays."""
    scenes, views, st = small
    g = st.g
    b = g.Batch(gpu_ctx, st.items)
    res = []
    for i, (si, v) in enumerate(views):
        lighting = S.default_lighting(v)
        res.append((v.object_to_clip, g.GLIMP_RESOLVE_LIT if i % 2 == 0 else g.GLIMP_RESOLVE_ALBEDO, lighting))
    b.set_resolve(res)
    b.launch_per_view()
    want = [fb.read_color() for fb in st.fbs]
    assert any((w != want[0][0, 0]).any() for w in want)
    for r in range(3):
        b.launch(graph=True)
        got = [fb.read_color() for fb in st.fbs]
        for i, (w, gt) in enumerate(zip(want, got)):
            assert (w == gt).all(), f"replay {r}: view {i} colour differs at {int((w != gt).sum())} pixels"
        for (ds, fb, v), (o2c, mode, lighting) in list(zip(st.items, res))[:3]:
            fb.clear(0x123, 0.0)
            fb.draw(ds, v.object_to_clip, v.meshlet_count // 2, 0)
            fb.draw(ds, v.object_to_clip, v.meshlet_count - v.meshlet_count // 2, v.meshlet_count // 2)
            fb.resolve(ds, o2c, mode, lighting)
    b.launch(graph=False)
    got = [fb.read_color() for fb in st.fbs]
    assert all((w == gt).all() for w, gt in zip(want, got))
    b.set_resolve(None)
    b.close()


@pytest.mark.parametrize("shape", [(32, 2048, 256), (32, 1024, 512), (32, 256, 1024), (16, 1024, 128),
                                   (16, 256, 256), (8, 512, 128), (64, 1024, 256), (64, 4096, 1024)])
def test_single_view_lit_batch_equals_separate_calls(gpu_ctx, shape):
    """The bench's frame: a batch of one view with its lit resolve gives the
    colour (light sprites included) and the vis-buffer of clear + draw +
    resolve as separate calls, eager and on every graph replay, for every bin
    size and block shape, including bins whose records overflow (capacity
    256)."""
    import glimpsw_amd as g

    scene, view = S.frame_1080p(num_meshlets=2048, width=960, height=540, tex_size=256, num_materials=4)
    ds = g.Scene(gpu_ctx, scene)
    fb = g.Framebuffer(gpu_ctx, view.width, view.height)
    ref = g.Framebuffer(gpu_ctx, view.width, view.height)
    lighting = S.default_lighting(view)
    planes = view.frustum_planes()
    ref.clear(0, 0.0)
    ref.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes)
    ref.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)
    want_c = ref.read_color()
    want_ids, want_d = ref.read()
    assert (want_ids != 0).any()
    bs, cap, th = shape
    b = g.Batch(gpu_ctx, [(ds, fb, view, planes, 0)], shape={"bin_size": bs, "bin_capacity": cap,
                                                              "raster_threads": th})
    b.set_resolve([(view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)])
    print(f"[frame batch] shape {shape}", flush=True)
    for r, graph in enumerate([False, True, True, True, False]):
        b.launch(graph=graph)
        got_c = fb.read_color()
        ids, d = fb.read()
        bad = got_c != want_c
        assert not bad.any(), f"launch {r} (graph={graph}): colour differs at {int(bad.sum())} pixels"
        assert (ids == want_ids).all() and (d.view(np.uint32) == want_d.view(np.uint32)).all(), \
            f"launch {r}: vis-buffer differs"
    b.close()
    ref.close()
    fb.close()
    ds.close()
