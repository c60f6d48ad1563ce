"""glTF import (Scene::ImportGltf, Scene.cpp:156-368) into meshlets.

The meshlet partition is this library's own (meshoptimizer, which the
reference uses, is not in the image), so it is checked by what any valid
partition must satisfy: every input triangle appears in exactly one meshlet
(as the same three positions, same winding), limits of 64 vertices and 128
triangles, no more meshlets than the bound, bounding spheres that contain
their vertices. Attribute packing is checked against its definition:
positions bit-exact, UVs as IEEE half (round to nearest even), normals and
tangents as octahedral bytes that decode to the input within the byte
precision and equal a float emulation of the reference's optimised
MapOctahedron (VRCP14PS reciprocal, one FMA for c * 0.5 + 0.5) with glm's
packUnorm4x8, tangent handedness from the sign of w; node matrices and TRS
compose down the hierarchy.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle
from glimpsw_amd import gltf
from glimpsw_amd import scene as S
from tests import scenes as T

SPONZA = "/root/reference/assets/models/Sponza/Sponza_LowPoly.gltf"


def _canon(tris):
    """Rotation-invariant key of each triangle (same winding): the smallest
    of its three corner rotations."""
    out = []
    for r in np.asarray(tris, np.float32).reshape(-1, 9).tolist():
        out.append(min(tuple(r[3 * k:] + r[:3 * k]) for k in range(3)))
    return out


def _meshlet_triangles(sc):
    out = []
    for m in range(sc.count):
        for t in range(int(sc.num_triangles[m])):
            out.append(sc.positions[m][:, sc.indices[m, :, t]].T.reshape(9))
    return np.array(out, np.float32)


def _check_partition(sc, pos, idx):
    assert (sc.num_vertices <= 64).all() and (sc.num_triangles <= 128).all()
    assert (sc.num_triangles > 0).all()
    got = sorted(_canon(_meshlet_triangles(sc)))
    want = sorted(_canon(pos[idx.reshape(-1, 3)].reshape(-1, 9)))
    assert got == want
    for m in range(sc.count):
        nv = int(sc.num_vertices[m])
        P = sc.positions[m][:, :nv].T.astype(np.float64)
        c, r = sc.bounds[m, :3].astype(np.float64), float(sc.bounds[m, 3])
        assert (np.linalg.norm(P - c, axis=1) <= r).all()
        assert int(sc.indices[m, :, : sc.num_triangles[m]].max()) < nv


def _unmap(b0, b1):
    u, v = b0 / 255.0 * 2 - 1, b1 / 255.0 * 2 - 1
    z = 1 - np.abs(u) - np.abs(v)
    t = np.maximum(-z, 0)
    x = u - np.where(u < 0, -t, t)
    y = v - np.where(v < 0, -t, t)
    n = np.stack([x, y, z], -1)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def test_torus_round_trip(tmp_path):
    pos, nrm, tan, uv, tris = T.torus_mesh()
    M = np.eye(4)
    M[:3, 3] = [0.5, -0.25, -4.0]
    path = str(tmp_path / "torus.gltf")
    gltf.write_gltf(path, pos, tris, nrm, tan, uv, matrix=M,
                    material={"doubleSided": True, "alphaMode": "MASK", "alphaCutoff": 0.3})
    model = gltf.import_gltf(path)
    sc = model.scene
    from glimpsw_amd import load
    bound = load().glimp_build_meshlets_bound(tris.size, 64, 128)
    assert sc.count <= bound and sc.count >= tris.shape[0] / 128
    _check_partition(sc, pos, tris)
    assert sc.materials == [(-1, True, int(np.float32(0.3) * np.float32(255) + np.float32(0.5)))]
    assert (sc.material_id == 0).all()
    assert len(model.nodes) == 1 and np.allclose(model.nodes[0].transform, M)
    # attributes: find each meshlet vertex in the input by position
    lookup = {tuple(p): i for i, p in enumerate(pos.tolist())}
    for m in range(0, sc.count, 7):
        for k in range(int(sc.num_vertices[m])):
            i = lookup[tuple(sc.positions[m][:, k].tolist())]
            tc = int(sc.texcoords[m, k])
            want = uv[i].astype(np.float16).view(np.uint16)
            assert tc & 0xFFFF == int(want[0]) and tc >> 16 == int(want[1])
            w = int(sc.normal_tangents[m, k])
            b = [(w >> (8 * q)) & 255 for q in range(4)]
            assert np.abs(_unmap(b[0], b[1]) - nrm[i]).max() < 0.02
            assert np.abs(_unmap(b[2], b[3]) - tan[i, :3]).max() < 0.02
            assert ((int(sc.tangent_handedness[m]) >> k) & 1) == int(tan[i, 3] < 0)


def test_normal_packing_matches_emulation():
    """glimp_pack_normal_tangents against a float32 emulation of the
    reference's MapOctahedron (VRCP14PS, FMAs) and glm::packUnorm4x8."""
    from glimpsw_amd import load

    rng = np.random.default_rng(3)
    n = rng.normal(size=(2048, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True).astype(np.float32)
    t = np.concatenate([n[::-1], np.ones((2048, 1), np.float32)], axis=1)
    out = np.zeros(2048, np.uint32)
    f32p = C.POINTER(C.c_float)
    load().glimp_pack_normal_tangents(n.ctypes.data_as(f32p), np.ascontiguousarray(t).ctypes.data_as(f32p), 2048,
                                      out.ctypes.data_as(C.POINTER(C.c_uint32)))
    L = oracle.lib()

    def f(x):
        return np.float32(x)

    def fma(a, b, c):   # single rounding: exact in float64 for float32 operands here
        return f(np.float64(a) * np.float64(b) + np.float64(c))

    def oct_bytes(v):
        x, y, z = (f(c) for c in v)
        w = f(L.orc_rcp14(float(f(f(abs(x)) + f(abs(y))) + f(abs(z)))))
        mz = f(-z * w)
        tt = mz if mz > 0 else f(0)
        sx = -tt if np.signbit(x) else tt
        sy = -tt if np.signbit(y) else tt
        u = fma(fma(x, w, sx), 0.5, 0.5)
        vv = fma(fma(y, w, sy), 0.5, 0.5)
        return [int(np.floor(f(min(max(c, 0.0), 1.0)) * f(255) + 0.5)) for c in (u, vv)]

    for i in range(0, 2048, 3):
        want = oct_bytes(n[i]) + oct_bytes(t[i, :3])
        got = [(int(out[i]) >> (8 * q)) & 255 for q in range(4)]
        assert got == want, (i, got, want)


def test_node_hierarchy(tmp_path):
    """Children compose with their parents: matrix and TRS nodes."""
    pos, nrm, tan, uv, tris = T.torus_mesh(nu=12, nv=6)
    path = str(tmp_path / "h.gltf")
    gltf.write_gltf(path, pos, tris)
    with open(path) as fh:
        doc = json.load(fh)
    doc["nodes"] = [{"children": [1], "translation": [1, 2, 3], "scale": [2, 2, 2]},
                    {"mesh": 0, "rotation": [0, np.sin(np.pi / 4), 0, np.cos(np.pi / 4)]}]
    with open(path, "w") as fh:
        json.dump(doc, fh)
    model = gltf.import_gltf(path)
    M = model.nodes[0].transform
    assert np.allclose(M @ [1, 0, 0, 1], [1, 2, 3 - 2, 1])       # y-rotation by 90 deg, then scale and move


@pytest.mark.skipif(not os.path.exists(SPONZA), reason="reference assets not present")
def test_sponza_low_poly():
    model = gltf.import_gltf(SPONZA)
    sc = model.scene
    assert sc.triangles() == 63084
    assert len(model.nodes) == 2 and np.allclose(model.nodes[0].transform[0, 0], 0.008)
    doc = gltf._Doc(SPONZA)
    for mi, (off, cnt) in enumerate(model.mesh_ranges):
        prim = doc.j["meshes"][mi]["primitives"][0]
        pos = doc.accessor(prim["attributes"]["POSITION"]).astype(np.float32)
        idx = doc.accessor(prim["indices"]).reshape(-1).astype(np.uint32)
        sub = S.Scene(sc.bounds[off:off + cnt], sc.num_vertices[off:off + cnt], sc.num_triangles[off:off + cnt],
                      sc.material_id[off:off + cnt], sc.tangent_handedness[off:off + cnt],
                      sc.positions[off:off + cnt], sc.texcoords[off:off + cnt], sc.normal_tangents[off:off + cnt],
                      sc.indices[off:off + cnt], [], [])
        _check_partition(sub, pos, idx)
