"""The oracle frame split over host threads by row bands (the CPU baseline's
all-core mode) equals the one-thread frame bit for bit: ids, depth and the
lit colour, light sprites included, for band counts that do and do not divide
the frame's tile rows."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S


@pytest.mark.parametrize("parts", [1, 3, 8, 37])
def test_row_bands_match_one_thread(parts):
    sc, view = S.checker_scene(num_meshlets=256, width=320, height=244)
    b = oracle.SceneBinding(sc)
    lighting = S.default_lighting(view)
    lighting.light_sprites = True
    ofb, _, _ = oracle.render_frame(sc, view, b)
    ids, depth = ofb.ids(), ofb.depth()
    rgba = oracle.resolve_lit(ofb, b, view.object_to_clip, lighting)
    bands = oracle.row_bands(view.height, parts)
    assert bands[0][0] == 0 and bands[-1][1] == view.height
    with ThreadPoolExecutor(4) as pool:
        tfb = oracle.render_frame_threads(sc, view, b, lighting, pool, bands)
    assert (tfb.pixels(0) == rgba).all()
    assert (tfb.depth().view(np.uint32) == depth.view(np.uint32)).all()
    assert (depth > 0).sum() > 1000
    # the ids of the banded draw, before its resolve
    fb = oracle.OracleFramebuffer(view.width, view.height)
    fb.clear(0, 0.0)
    bitmap, _ = oracle.cull(b, view.frustum_planes(), 0, view.meshlet_count)
    for r in bands:
        oracle.draw(fb, b, view.object_to_clip, view.meshlet_count, 0, bitmap, rows=r)
    assert (fb.ids() == ids).all()
