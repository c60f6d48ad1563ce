"""glimp_draw_meshlets_static, the draw that allocates nothing (its scratch is
a static device array), against glimp_draw_meshlets on the direct route:
two calls in a row on one framebuffer pair, the second with fewer meshlets
and another view, must give bit-identical vis-buffers and counters, and a
call over the 1024-meshlet limit must be refused without drawing.

The calls go through the C ABI (ctypes) of libglimp.so."""
import ctypes as C

import numpy as np
import pytest

from glimpsw_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu

GLIMP_E_LIMIT = -4          # include/glimp.h


def _desc(g, view, count, offset, planes):
    d = g.DrawDesc()
    d.meshlet_offset = offset
    d.meshlet_count = count
    for i in range(16):
        d.object_to_clip[i] = float(view.object_to_clip[i])
    keep = None
    if planes is not None:
        keep = np.ascontiguousarray(planes, np.float32)
        d.frustum_planes = keep.ctypes.data_as(C.POINTER(C.c_float))
    d.flags = g.GLIMP_DRAW_DIRECT
    d.pipeline = g.GLIMP_PIPELINE_VISBUFFER
    return d, keep


def _frames(g, ctx, ds, fb_static, fb_plain, view, count, offset):
    planes = view.frustum_planes()
    d, keep = _desc(g, view, count, offset, planes)
    out = []
    for fb, fn in ((fb_static, ctx.lib.glimp_draw_meshlets_static), (fb_plain, ctx.lib.glimp_draw_meshlets)):
        fb.clear(0x0BADF00D, 0.0)
        ctx.reset_counters()
        rc = fn(ctx.handle, fb.handle, ds.handle, C.byref(d), None)
        assert rc == 0, ctx.lib.glimp_last_error()
        ids, depth = fb.read()
        out.append((ids, depth, ctx.counters()))
    del keep
    return out


def test_static_draw_matches_draw(gpu_ctx):
    import glimpsw_amd as g

    ctx = gpu_ctx
    # alpha-tested cutouts and near-plane crossers: every scratch region in use
    sc = T.alpha_scene(1000, seed_offset=4, spread=6.0, near_frac=0.15)
    ds = g.Scene(ctx, sc)
    fb_s = g.Framebuffer(ctx, 320, 240)
    fb_p = g.Framebuffer(ctx, 320, 240)
    try:
        view1 = S.make_view(320, 240, count=sc.count)
        view2 = S.make_view(320, 240, 75.0, S.view_matrix((0.5, -0.3, 1.0), 0.4, -0.2), 0, sc.count)
        for label, view, count, offset in (("1000 meshlets", view1, 1000, 0), ("300 meshlets", view2, 300, 650)):
            (si, sd, sc_), (pi, pd, pc) = _frames(g, ctx, ds, fb_s, fb_p, view, count, offset)
            print(f"[no-alloc] {label}: {int((pd > 0).sum())} covered pixels", flush=True)
            assert (pd > 0).sum() > 1000
            assert (si == pi).all() and (sd.view(np.uint32) == pd.view(np.uint32)).all(), label
            for k in ("triangles_processed", "triangles_rasterized", "triangles_clipped", "fragments_covered",
                      "dropped"):
                assert sc_[k] == pc[k], (label, k, sc_[k], pc[k])
        # over the limit: refused, the framebuffer untouched
        big = T.small_scene(1100, seed_offset=9, spread=6.0)
        dbig = g.Scene(ctx, big)
        vbig = S.make_view(320, 240, count=big.count)
        d, keep = _desc(g, vbig, 1100, 0, vbig.frustum_planes())
        fb_s.clear(0x0BADF00D, 0.0)
        rc = ctx.lib.glimp_draw_meshlets_static(ctx.handle, fb_s.handle, dbig.handle, C.byref(d), None)
        assert rc == GLIMP_E_LIMIT and b"1024" in ctx.lib.glimp_last_error()
        ids, _ = fb_s.read()
        assert (ids == 0x0BADF00D).all()
        dbig.close()
    finally:
        fb_s.close()
        fb_p.close()
        ds.close()
