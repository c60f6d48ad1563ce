"""A glTF model through the path: a torus written as .gltf, imported into
meshlets (glimpsw_amd.gltf), drawn per node with its model matrix
(ObjectToClipMat = proj * view * model, Shading.h:35-39) and resolved; ids,
depth and albedo bit-exact with the oracle, lit within 1/255."""
import dataclasses

import numpy as np
import pytest

import oracle
from glimpsw_amd import gltf
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T

pytestmark = pytest.mark.gpu


def test_imported_torus(gpu_ctx, tmp_path):
    import glimpsw_amd as g

    pos, nrm, tan, uv, tris = T.torus_mesh(nu=160, nv=64)
    M = np.eye(4)
    c, s_ = np.cos(0.6), np.sin(0.6)
    M[:3, :3] = [[1, 0, 0], [0, c, -s_], [0, s_, c]]
    M[:3, 3] = [0.2, -0.1, -3.0]
    path = str(tmp_path / "torus.gltf")
    gltf.write_gltf(path, pos, tris, nrm, tan, uv, matrix=M, material={"doubleSided": False})
    model = gltf.import_gltf(path)
    sc = dataclasses.replace(model.scene, materials=[(0, False, 255)], textures=[textures.checker_texture(256)])
    view = S.make_view(640, 480, count=sc.count)
    ((off, cnt, model_mat),) = list(model.draws())
    P = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    o2c = (P @ model_mat.astype(np.float64).reshape(4, 4).T).T.reshape(16).astype(np.float32)
    print(f"[gltf] torus: {sc.count} meshlets, {sc.triangles()} triangles", flush=True)

    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, view.width, view.height)
    fb.clear(0, 0.0)
    fb.draw(ds, o2c, cnt, off)
    ids, depth = fb.read()
    fb.resolve(ds, o2c, g.GLIMP_RESOLVE_ALBEDO)
    alb = fb.read_color()
    fb.close()
    ds.close()

    b = oracle.SceneBinding(sc)
    ofb = oracle.OracleFramebuffer(view.width, view.height)
    ofb.clear(0, 0.0)
    oracle.draw(ofb, b, o2c, cnt, off)
    assert (depth > 0).sum() > 20000
    assert (ids == ofb.ids()).all() and (depth.view(np.uint32) == ofb.depth().view(np.uint32)).all()
    assert (alb == oracle.resolve_albedo(ofb, b, o2c, off)).all()
