"""The C++ swr host layer (include/glimp_swr.hpp) end to end: a C++ program
(tests/cpp/swr_host.cpp) renders a seeded scene through
Rasterizer::DrawMeshlets / ShadingContext::ResolveDebug on the GPU; its ids,
depth and albedo must equal the oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "swr_host")


def _write_scene(d, sc, view):
    t = sc.textures[0]
    planes = view.frustum_planes()
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{sc.count} {view.width} {view.height} {t.width} {t.height} {t.mip_levels} {t.num_layers} "
                f"{t.row_shift} {t.layer_stride} {t.data.size}\n")
        f.write(" ".join(str(int(v)) for v in t.mip_offsets) + "\n")
        f.write(" ".join(repr(float(v)) for v in view.object_to_clip) + "\n")
        f.write(" ".join(repr(float(v)) for v in planes) + "\n")
    for name in ("bounds", "num_vertices", "num_triangles", "material_id", "tangent_handedness", "positions",
                 "texcoords", "normal_tangents", "indices"):
        np.ascontiguousarray(getattr(sc, name)).tofile(os.path.join(d, name + ".bin"))
    t.data.tofile(os.path.join(d, "texels.bin"))


def test_swr_host_program(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/swr_host is not built (run __graft_entry__.build())")
    sc, view = S.checker_scene(512, 320, 240)
    _write_scene(str(tmp_path), sc, view)
    print("[swr_host] checker scene 512 meshlets 320x240", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = view.width * view.height
    ids = np.fromfile(tmp_path / "ids.bin", np.uint32).reshape(view.height, view.width)
    depth = np.fromfile(tmp_path / "depth.bin", np.uint32).reshape(view.height, view.width)
    color = np.fromfile(tmp_path / "color.bin", np.uint32).reshape(view.height, view.width)
    bitmap = np.fromfile(tmp_path / "bitmap.bin", np.uint16)
    assert ids.size == n
    ofb, _, b = oracle.render_frame(sc, view)
    assert (ids == ofb.ids()).all()
    assert (depth == ofb.depth().view(np.uint32)).all()
    want_bitmap, visible = oracle.cull(b, view.frustum_planes(), 0, sc.count)
    assert (bitmap == want_bitmap).all()
    assert f"{visible} visible" in r.stdout
    assert (color == oracle.resolve_albedo(ofb, b, view.object_to_clip)).all()
    # the overdraw table over the same frame
    bm, _ = oracle.cull(b, view.frustum_planes(), 0, sc.count)
    od = oracle.OracleFramebuffer(view.width, view.height)
    od.clear(0, 0.0)
    oracle.draw(od, b, view.object_to_clip, sc.count, 0, bm, pipeline=oracle.PIPELINE_OVERDRAW)
    words = np.fromfile(tmp_path / "od_words.bin", np.uint32).reshape(view.height, view.width)
    odepth = np.fromfile(tmp_path / "od_depth.bin", np.uint32).reshape(view.height, view.width)
    ocolor = np.fromfile(tmp_path / "od_color.bin", np.uint32).reshape(view.height, view.width)
    assert (words == od.pixels(0)).all() and (odepth == od.pixels(1)).all()
    assert (ocolor == oracle.resolve_debug(od, b, view.object_to_clip, oracle.DEBUG_OVERDRAW_QUAD)).all()
