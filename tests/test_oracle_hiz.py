"""Pins the oracle's occlusion culling (CullMeshlets with a depth map,
Shading.cpp:775-869, texutil::DownsampleDepth, ImageHelpers.cpp:149-247)
against what the geometry fixes.

ProjectSphere and the per-meshlet test arithmetic are pinned bit for bit
against the reference's compiled form (test_oracle_fastmath.py). Here:
 - every pyramid texel is the minimum depth of its 2^(m+1) pixel block,
   computed independently with numpy;
 - with the depth map of the same frame and the same camera, occlusion
   culling is conservative: the frame drawn from the occlusion-culled
   bitmap equals the frame drawn without it, pixel for pixel, while the
   wall in front of the scene removes a good share of the meshlets;
 - a meshlet whose depth-map footprint shows only the wall and whose
   sphere lies behind the wall is culled; one whose sphere lies in front of
   everything under its footprint is kept.
"""
import numpy as np

import oracle
from glimpsw_amd import scene as S
from tests import scenes as T

W, H = 640, 480


def _frame():
    sc = T.occluder_scene()
    view = S.make_view(W, H, count=sc.count)
    ofb, _, b = oracle.render_frame(sc, view)
    return sc, view, ofb, b


def test_pyramid_is_block_minimum():
    _, _, ofb, _ = _frame()
    depth = ofb.depth()
    hiz = oracle.hiz_build(ofb)
    off = 0
    for m, (w, h) in enumerate(oracle.hiz_level_dims(W, H)):
        s = 1 << (m + 1)
        pad = np.full((h * s, w * s), np.finfo(np.float32).max, np.float32)
        pad[:H, :W] = depth[: h * s, : w * s]
        want = pad.reshape(h, s, w, s).min(axis=(1, 3))
        got = hiz[off: off + w * h].reshape(h, w)
        assert (got == want).all(), m
        off += w * h
    assert off == hiz.size


def test_occlusion_is_conservative_and_culls():
    sc, view, ofb, b = _frame()
    ids0, depth0 = ofb.ids().copy(), ofb.depth().copy()
    keep = []
    o2pv, scale, proj = S.occlusion_terms(view.proj, view.view)
    occ = oracle.occlusion(oracle.hiz_build(ofb), W, H, o2pv, scale, proj, keep)
    bm, n_occ = oracle.cull_occlusion(b, view.frustum_planes(), 0, sc.count, occ)
    _, n_frustum = oracle.cull(b, view.frustum_planes(), 0, sc.count)
    assert n_occ < 0.8 * n_frustum
    fb = oracle.OracleFramebuffer(W, H)
    fb.clear(0, 0.0)
    oracle.draw(fb, b, view.object_to_clip, sc.count, 0, bm)
    assert (fb.ids() == ids0).all() and (fb.depth().view(np.uint32) == depth0.view(np.uint32)).all()


def test_spheres_behind_the_wall_are_culled():
    """A sphere whose depth-map footprint (the texel range the test reads,
    from the probe-pinned box) shows only the wall, and whose nearest point
    lies clearly behind the wall, is culled; a sphere whose nearest point
    lies in front of everything under its footprint is kept."""
    import ctypes as C

    sc, view, ofb, b = _frame()
    depth = ofb.depth()
    keep = []
    o2pv, scale, proj = S.occlusion_terms(view.proj, view.view)
    occ = oracle.occlusion(oracle.hiz_build(ofb), W, H, o2pv, scale, proj, keep)
    bm, _ = oracle.cull_occlusion(b, view.frustum_planes(), 0, sc.count, occ)
    bf, _ = oracle.cull(b, view.frustum_planes(), 0, sc.count)
    kept = np.unpackbits(bm.view(np.uint8), bitorder="little")[: sc.count].astype(bool)
    in_frustum = np.unpackbits(bf.view(np.uint8), bitorder="little")[: sc.count].astype(bool)
    wall_depth = float(view.proj[14]) / 2.5
    behind = in_front = 0
    for g0 in range(0, sc.count, 16):
        n = min(16, sc.count - g0)
        sph = np.zeros((4, 16), np.float32)
        sph[:, :n] = sc.bounds[g0:g0 + n].T
        mask = np.zeros(16, np.int32); bb = np.zeros(64, np.float32); mip = np.zeros(16, np.int32)
        rect = np.zeros(64, np.int32); dsph = np.zeros(16, np.float32)
        f32p, i32p = C.POINTER(C.c_float), C.POINTER(C.c_int32)
        oracle.lib().orc_hiz_sphere16(sph.ctypes.data_as(f32p), C.byref(occ), mask.ctypes.data_as(i32p),
                                      bb.ctypes.data_as(f32p), mip.ctypes.data_as(i32p), rect.ctypes.data_as(i32p),
                                      dsph.ctypes.data_as(f32p))
        for l in range(n):
            i = g0 + l
            if i == sc.count - 1 or not in_frustum[i] or not mask[l]:
                continue
            lv = int(mip[l])
            x0, y0, x1, y1 = (int(rect[k * 16 + l]) for k in range(4))
            if x0 > x1 or y0 > y1:
                continue
            under = depth[y0 << lv: min((y1 + 1) << lv, H), x0 << lv: min((x1 + 1) << lv, W)]
            if np.allclose(under, wall_depth, rtol=1e-5) and dsph[l] < 0.95 * wall_depth:
                behind += 1
                assert not kept[i], i
            if dsph[l] > 1.05 * under.max():
                in_front += 1
                assert kept[i], i
    assert behind > 10 and in_front > 10, (behind, in_front)
