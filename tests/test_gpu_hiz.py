"""Occlusion culling on the GPU (CullMeshlets with the previous frame's depth
map, Shading.cpp:775-869; texutil::DownsampleDepth, ImageHelpers.cpp:241-245):
the depth pyramid, the culled bitmap and the frame drawn with occlusion are
bit-exact with the oracle."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu

W, H = 640, 480


def _setup(gpu_ctx, view_of=None):
    import glimpsw_amd as g

    sc = T.occluder_scene()
    view = S.make_view(W, H, count=sc.count)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.render_frame(gpu_ctx, ds, view)
    dm = g.DepthMap(gpu_ctx, W, H)
    dm.build(fb)
    ofb, _, b = oracle.render_frame(sc, view)
    return g, sc, view, ds, fb, dm, ofb, b


def test_depth_pyramid(gpu_ctx):
    g, sc, view, ds, fb, dm, ofb, b = _setup(gpu_ctx)
    got, levels = dm.read()
    want = oracle.hiz_build(ofb)
    print(f"[hiz] {levels} levels, {got.size} texels", flush=True)
    assert levels == oracle.hiz_levels(W, H) and got.size == want.size
    assert (got.view(np.uint32) == want.view(np.uint32)).all()
    for o in (dm, fb, ds):
        o.close()


@pytest.mark.parametrize("yaw", [0.0, 0.08])
def test_occlusion_cull_and_draw(gpu_ctx, yaw):
    """The depth map of one frame culls the next frame's meshlets (same
    camera, or turned by `yaw`); the bitmap and the frame match the oracle."""
    g, sc, view, ds, fb, dm, ofb, b = _setup(gpu_ctx)
    cur = S.make_view(W, H, view=S.view_matrix((0.0, 0.0, 0.0), yaw, 0.0), count=sc.count)
    o2pv, scale, proj = S.occlusion_terms(view.proj, view.view)    # the depth map's camera
    occ = g.Occlusion(dm, o2pv, scale, proj)
    bits, n = ds.cull_occlusion(cur.frustum_planes(), 0, sc.count, occ)
    keep = []
    oocc = oracle.occlusion(oracle.hiz_build(ofb), W, H, o2pv, scale, proj, keep)
    obits, on = oracle.cull_occlusion(b, cur.frustum_planes(), 0, sc.count, oocc)
    print(f"[hiz] yaw {yaw}: visible gpu {n} oracle {on}", flush=True)
    assert n == on and (bits == obits).all()
    fb2 = g.Framebuffer(gpu_ctx, W, H)
    fb2.clear(0, 0.0)
    fb2.draw(ds, cur.object_to_clip, sc.count, 0, cur.frustum_planes(), occlusion=occ)
    gi, gd = fb2.read()
    of2 = oracle.OracleFramebuffer(W, H)
    of2.clear(0, 0.0)
    oracle.draw(of2, b, cur.object_to_clip, sc.count, 0, obits)
    assert (gi == of2.ids()).all() and (gd.view(np.uint32) == of2.depth().view(np.uint32)).all()
    for o in (fb2, dm, fb, ds):
        o.close()


def test_occlusion_in_a_batch(gpu_ctx):
    """Two views of the occluder scene, each culled against its own previous
    frame's depth map inside one batch: graph replay, eager launch and
    per-view draws agree with each other and with the oracle."""
    import glimpsw_amd as g

    sc = T.occluder_scene()
    ds = g.Scene(gpu_ctx, sc)
    cams = [S.make_view(W, H, count=sc.count),
            S.make_view(W, H, view=S.view_matrix((0.0, 0.0, 0.0), 0.08, 0.0), count=sc.count)]
    prev = [g.render_frame(gpu_ctx, ds, v) for v in cams]
    dms = [g.DepthMap(gpu_ctx, W, H) for _ in cams]
    for dm, fb in zip(dms, prev):
        dm.build(fb)
    cur = [S.make_view(W, H, view=S.view_matrix((0.0, 0.0, 0.0), yaw, 0.0), count=sc.count) for yaw in (0.03, 0.11)]
    occs = []
    for v, dm in zip(cams, dms):
        o2pv, scale, proj = S.occlusion_terms(v.proj, v.view)
        occs.append(g.Occlusion(dm, o2pv, scale, proj))
    fbs = [g.Framebuffer(gpu_ctx, W, H) for _ in cur]
    batch = g.Batch(gpu_ctx, [(ds, fb, v, v.frustum_planes(), 0, o) for fb, v, o in zip(fbs, cur, occs)])
    try:
        want = []
        for v, cam, f in zip(cur, cams, prev):
            ofb, _, b = oracle.render_frame(sc, cam)
            o2pv, scale, proj = S.occlusion_terms(cam.proj, cam.view)
            keep = []
            oocc = oracle.occlusion(oracle.hiz_build(ofb), W, H, o2pv, scale, proj, keep)
            bits, n = oracle.cull_occlusion(b, v.frustum_planes(), 0, sc.count, oocc)
            _, nf = oracle.cull(b, v.frustum_planes(), 0, sc.count)
            assert n < nf
            o2 = oracle.OracleFramebuffer(W, H)
            o2.clear(0, 0.0)
            oracle.draw(o2, b, v.object_to_clip, sc.count, 0, bits)
            want.append((o2.ids(), o2.depth()))
        for mode in ("graph", "eager", "graph", "per_view"):
            if mode == "per_view":
                batch.launch_per_view()
            else:
                batch.launch(graph=(mode == "graph"))
            for i, fb in enumerate(fbs):
                ids, depth = fb.read()
                print(f"[hiz] batch {mode} view {i}", flush=True)
                assert (ids == want[i][0]).all() and (depth.view(np.uint32) == want[i][1].view(np.uint32)).all()
    finally:
        batch.close()
        for o in fbs + prev + dms:
            o.close()
        ds.close()
