"""glTF 2.0 import into the meshlet scene format: Scene::ImportGltf
(Scene.cpp:156-368) without its texture loading.

Reads the JSON document and its binary buffers (files next to it or base64
data URIs), then for every mesh primitive: indices, POSITION, TEXCOORD_0,
NORMAL and TANGENT; builds meshlets of at most 64 vertices and 128
triangles with ``glimp_build_meshlets`` (the reference calls meshoptimizer's
meshopt_buildMeshlets with the same limits, Scene.cpp:214-217; that library
is absent here, so partitions differ) and bounding spheres with
``glimp_meshlet_bounds``; converts the vertex attributes as the reference
does (Scene.cpp:244-280): positions as they are, texture coordinates as
packHalf2x16 (round to nearest even), normal and tangent as two octahedral
byte pairs (``glimp_pack_normal_tangents``), the tangent's sign bit into
TangentHandedness. Materials keep double-sidedness and the alpha cutoff
(``mask`` mode: byte(cutoff * 255 + 0.5), else 255, Scene.cpp:179-183); their
images are not decoded (no image codecs here), so materials carry no texture
unless the caller assigns one. Node transforms are composed down the
hierarchy (matrix, or translation * rotation * scale) into each node's
global transform, the model matrix of its draws.
"""
from __future__ import annotations

import base64
import ctypes as C
import dataclasses
import json
import os

import numpy as np

from . import scene as S

_DTYPE = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_NORM = {5120: 127.0, 5121: 255.0, 5122: 32767.0, 5123: 65535.0}

MAX_VERTICES, MAX_TRIANGLES = 64, 128


@dataclasses.dataclass
class ModelNode:
    mesh: int                      # index into Model.mesh_ranges
    transform: np.ndarray          # (4, 4) float64 global transform (column vectors)


@dataclasses.dataclass
class Model:
    scene: S.Scene                 # meshlets of all meshes, in mesh order
    mesh_ranges: list              # per mesh: (meshlet offset, meshlet count)
    nodes: list                    # ModelNode per node with a mesh, depth-first from the scene roots

    def draws(self):
        """(meshlet offset, count, column-major model matrix) per node."""
        for n in self.nodes:
            off, cnt = self.mesh_ranges[n.mesh]
            if cnt:
                yield off, cnt, n.transform.T.reshape(16).astype(np.float32)


class _Doc:
    def __init__(self, path: str):
        self.base = os.path.dirname(os.path.abspath(path))
        with open(path, "r", encoding="utf-8") as f:
            self.j = json.load(f)
        self.buffers = []
        for b in self.j.get("buffers", []):
            uri = b.get("uri")
            if uri is None:
                raise ValueError("glTF: binary chunks (.glb) are not supported")
            if uri.startswith("data:"):
                data = base64.b64decode(uri.split(",", 1)[1])
            else:
                with open(os.path.join(self.base, uri), "rb") as f:
                    data = f.read()
            if len(data) < b["byteLength"]:
                raise ValueError(f"glTF: buffer {uri[:40]} is shorter than its byteLength")
            self.buffers.append(np.frombuffer(data, np.uint8))

    def accessor(self, idx: int) -> np.ndarray:
        """(count, components) array; normalised integers become float32."""
        a = self.j["accessors"][idx]
        if "sparse" in a:
            raise ValueError("glTF: sparse accessors are not supported")
        dt = np.dtype(_DTYPE[a["componentType"]])
        nc = _NCOMP[a["type"]]
        count = a["count"]
        if "bufferView" not in a:
            out = np.zeros((count, nc), dt)
        else:
            bv = self.j["bufferViews"][a["bufferView"]]
            buf = self.buffers[bv["buffer"]]
            start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
            stride = bv.get("byteStride", 0) or dt.itemsize * nc
            rows = np.lib.stride_tricks.as_strided(buf[start:], shape=(count, dt.itemsize * nc),
                                                   strides=(stride, 1)) if count else np.zeros((0, dt.itemsize * nc), np.uint8)
            out = np.ascontiguousarray(rows).view(dt).reshape(count, nc)
        if a.get("normalized") and a["componentType"] in _NORM:
            out = np.maximum(out.astype(np.float32) / np.float32(_NORM[a["componentType"]]), -1.0).astype(np.float32)
        return out


def _node_matrix(n: dict) -> np.ndarray:
    if "matrix" in n:
        return np.asarray(n["matrix"], np.float64).reshape(4, 4).T
    T = np.eye(4)
    T[:3, 3] = n.get("translation", [0, 0, 0])
    x, y, z, w = n.get("rotation", [0, 0, 0, 1])
    R = np.eye(4)
    R[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    Sc = np.diag(list(n.get("scale", [1, 1, 1])) + [1.0])
    return T @ R @ Sc


def _build_meshlets(lib, idx: np.ndarray, pos: np.ndarray):
    u32p, f32p = C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    idx = np.ascontiguousarray(idx, np.uint32)
    pos = np.ascontiguousarray(pos, np.float32)
    bound = lib.glimp_build_meshlets_bound(idx.size, MAX_VERTICES, MAX_TRIANGLES)
    ranges = (_ranges_type() * max(bound, 1))()
    mverts = np.zeros(max(idx.size, 1), np.uint32)
    mtris = np.zeros(max(idx.size, 1), np.uint8)
    n = C.c_size_t()
    rc = lib.glimp_build_meshlets(idx.ctypes.data_as(u32p), idx.size, pos.ctypes.data_as(f32p), pos.shape[0], 12,
                                  MAX_VERTICES, MAX_TRIANGLES, ranges, mverts.ctypes.data_as(u32p),
                                  mtris.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n))
    if rc != 0:
        raise ValueError(f"glimp_build_meshlets failed ({rc})")
    return [(r.vertex_offset, r.triangle_offset, r.vertex_count, r.triangle_count) for r in ranges[:n.value]], \
        mverts, mtris


def _ranges_type():
    from . import MeshletRange
    return MeshletRange


def import_gltf(path: str) -> Model:
    """Loads a .gltf file into meshlets (see the module docstring)."""
    from . import load

    lib = load()
    doc = _Doc(path)
    j = doc.j
    materials = []
    for m in j.get("materials", []):
        cutoff = 255
        if m.get("alphaMode") == "MASK":
            cutoff = int(np.float32(m.get("alphaCutoff", 0.5)) * np.float32(255.0) + np.float32(0.5))
        materials.append((-1, bool(m.get("doubleSided", False)), cutoff))
    rows = {k: [] for k in ("bounds", "nv", "nt", "mat", "hand", "pos", "tc", "ntg", "idx")}
    mesh_ranges = []
    f32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    total = 0
    for mesh in j.get("meshes", []):
        start = total
        for prim in mesh["primitives"]:
            if prim.get("mode", 4) != 4:
                raise ValueError("glTF: only triangle lists are supported")
            attr = prim["attributes"]
            pos = doc.accessor(attr["POSITION"]).astype(np.float32)
            idx = (doc.accessor(prim["indices"]).reshape(-1).astype(np.uint32) if "indices" in prim
                   else np.arange(pos.shape[0], dtype=np.uint32))
            uv = doc.accessor(attr["TEXCOORD_0"]).astype(np.float32) if "TEXCOORD_0" in attr else None
            nrm = doc.accessor(attr["NORMAL"]).astype(np.float32) if "NORMAL" in attr else None
            tan = doc.accessor(attr["TANGENT"]).astype(np.float32) if "TANGENT" in attr else None
            mat = prim.get("material")
            ranges, mverts, mtris = _build_meshlets(lib, idx, pos)
            for vo, to, vc, tc in ranges:
                vids = mverts[vo:vo + vc]
                sphere = np.zeros(4, np.float32)
                lib.glimp_meshlet_bounds(np.ascontiguousarray(vids).ctypes.data_as(u32p), vc,
                                         pos.ctypes.data_as(f32p), 12, sphere.ctypes.data_as(f32p))
                P = np.zeros((3, 64), np.float32)
                P[:, :vc] = pos[vids].T
                T = np.zeros(64, np.uint32)
                if uv is not None:
                    hb = uv[vids].astype(np.float16).view(np.uint16).astype(np.uint32)
                    T[:vc] = hb[:, 0] | (hb[:, 1] << 16)
                NT = np.zeros(64, np.uint32)
                hand = 0
                if nrm is not None:
                    n = np.ascontiguousarray(nrm[vids])
                    t = np.ascontiguousarray(tan[vids]) if tan is not None else None
                    packed = np.zeros(vc, np.uint32)
                    lib.glimp_pack_normal_tangents(n.ctypes.data_as(f32p), t.ctypes.data_as(f32p) if t is not None
                                                   else None, vc, packed.ctypes.data_as(u32p))
                    NT[:vc] = packed
                    if t is not None:
                        for i in np.nonzero(t[:, 3] < 0)[0]:
                            hand |= 1 << int(i)
                I = np.zeros((3, 128), np.uint8)
                I[:, :tc] = mtris[to:to + 3 * tc].reshape(tc, 3).T
                rows["bounds"].append(sphere)
                rows["nv"].append(vc)
                rows["nt"].append(tc)
                rows["mat"].append(0xFFFFFFFF if mat is None else mat)
                rows["hand"].append(hand)
                rows["pos"].append(P)
                rows["tc"].append(T)
                rows["ntg"].append(NT)
                rows["idx"].append(I)
                total += 1
        mesh_ranges.append((start, total - start))
    sc = S.Scene(np.array(rows["bounds"], np.float32).reshape(-1, 4), np.array(rows["nv"], np.uint8),
                 np.array(rows["nt"], np.uint8), np.array(rows["mat"], np.uint32),
                 np.array(rows["hand"], np.uint64), np.array(rows["pos"], np.float32).reshape(-1, 3, 64),
                 np.array(rows["tc"], np.uint32).reshape(-1, 64), np.array(rows["ntg"], np.uint32).reshape(-1, 64),
                 np.array(rows["idx"], np.uint8).reshape(-1, 3, 128), materials, [])
    nodes = []
    scene_idx = j.get("scene", 0)
    roots = j["scenes"][scene_idx]["nodes"] if j.get("scenes") else list(range(len(j.get("nodes", []))))

    def walk(i, parent):
        n = j["nodes"][i]
        M = parent @ _node_matrix(n)
        if "mesh" in n:
            nodes.append(ModelNode(n["mesh"], M))
        for c in n.get("children", []):
            walk(c, M)

    for r in roots:
        walk(r, np.eye(4))
    return Model(sc, mesh_ranges, nodes)


def write_gltf(path: str, positions: np.ndarray, indices: np.ndarray, normals=None, tangents=None, texcoords=None,
               matrix=None, material=None) -> None:
    """Writes a one-mesh, one-node .gltf with an external .bin (test scenes)."""
    chunks, views, accessors = [], [], []
    offset = 0

    def add(arr, ctype, typ, target=None):
        nonlocal offset
        b = np.ascontiguousarray(arr).tobytes()
        pad = (-len(b)) % 4
        chunks.append(b + b"\0" * pad)
        v = {"buffer": 0, "byteOffset": offset, "byteLength": len(b)}
        if target:
            v["target"] = target
        views.append(v)
        acc = {"bufferView": len(views) - 1, "componentType": ctype, "count": int(arr.shape[0]), "type": typ}
        if typ == "VEC3" and ctype == 5126:
            acc["min"] = np.asarray(arr).min(axis=0).tolist()
            acc["max"] = np.asarray(arr).max(axis=0).tolist()
        accessors.append(acc)
        offset += len(b) + pad
        return len(accessors) - 1

    attrs = {"POSITION": add(positions.astype(np.float32), 5126, "VEC3", 34962)}
    if normals is not None:
        attrs["NORMAL"] = add(normals.astype(np.float32), 5126, "VEC3", 34962)
    if tangents is not None:
        attrs["TANGENT"] = add(tangents.astype(np.float32), 5126, "VEC4", 34962)
    if texcoords is not None:
        attrs["TEXCOORD_0"] = add(texcoords.astype(np.float32), 5126, "VEC2", 34962)
    ind = add(indices.astype(np.uint32).reshape(-1, 1), 5125, "SCALAR", 34963)
    prim = {"attributes": attrs, "indices": ind}
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}],
           "meshes": [{"primitives": [prim]}], "accessors": accessors, "bufferViews": views}
    if material is not None:
        prim["material"] = 0
        doc["materials"] = [material]
    node = {"mesh": 0}
    if matrix is not None:
        node["matrix"] = np.asarray(matrix, np.float64).T.reshape(16).tolist()
    doc["nodes"] = [node]
    binname = os.path.splitext(os.path.basename(path))[0] + ".bin"
    doc["buffers"] = [{"uri": binname, "byteLength": offset}]
    with open(os.path.join(os.path.dirname(os.path.abspath(path)), binname), "wb") as f:
        f.write(b"".join(chunks))
    with open(path, "w", encoding="utf-8") as f:
        json.dump(doc, f)
