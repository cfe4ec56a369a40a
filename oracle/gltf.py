"""CPU restatement of the reference's glTF scene loading -- test
infrastructure only (the reference results for xrh_scene_gltf; nothing in the product
imports it).

Follows AglTFModel::loadglTFFile / loadMaterials / loadNode
(src/Asset/Model.cpp:69-260) and the model concatenation of AScene::AScene
(src/Asset/Scene.cpp:16-50), with the matrix arithmetic of the vendored glm
0.9.9.8 in float32, one IEEE operation at a time (numpy float32 scalars):

* mat4 * vec4 (detail/type_mat4x4.inl:536-581): (m0 x + m1 y) + (m2 z + m3 w);
* mat4 * mat4 (:630-648): column i = ((A0 b0 + A1 b1) + A2 b2) + A3 b3;
* translate / scale (ext/matrix_transform.inl:10-15, 78-86);
* mat3_cast of a quaternion (gtc/quaternion.inl:41-66);
* inverse (detail/func_matrix.inl:294-351);
* normalize = v * (1 / sqrt((x x + y y) + z z)) (detail/func_geometric.inl:48-55, 82-89;
  detail/func_exponential.inl:136-139).

Parity pins beyond this restatement live in tests/test_cpu_gltf.py: closed
forms for rotations, TRS nodes, parent chains and inverse-transpose normals,
and the accessor min/max the cb2 file itself records.
"""
from __future__ import annotations

import base64
import json
import os
import urllib.parse

import numpy as np

F = np.float32
# AVertex (36 B) and AMaterial (32 B), field names as in xvk_ray_amd.native
VERTEX_DTYPE = np.dtype([("pos", "<f4", 3), ("normal", "<f4", 3), ("tex_coord", "<f4", 2),
                         ("material_index", "<i4")])
MATERIAL_DTYPE = np.dtype([("base_color", "<f4", 4), ("base_color_texture_index", "<i4"), ("fuzziness", "<f4"),
                           ("refraction_index", "<f4"), ("material_model", "<u4")])
LAMBERTIAN, DIFFUSE_LIGHT = 0, 4


def identity():
    return [[F(1) if c == k else F(0) for k in range(4)] for c in range(4)]


def mul_mv(m, v):
    out = []
    for k in range(4):
        add0 = m[0][k] * v[0] + m[1][k] * v[1]
        add1 = m[2][k] * v[2] + m[3][k] * v[3]
        out.append(add0 + add1)
    return out


def mul_mm(a, b):
    return [[((a[0][k] * b[i][0] + a[1][k] * b[i][1]) + a[2][k] * b[i][2]) + a[3][k] * b[i][3] for k in range(4)]
            for i in range(4)]


def translate(m, t):
    r = [col[:] for col in m]
    r[3] = [((m[0][k] * t[0] + m[1][k] * t[1]) + m[2][k] * t[2]) + m[3][k] for k in range(4)]
    return r


def scale(m, s):
    r = [col[:] for col in m]
    for c in range(3):
        r[c] = [m[c][k] * s[c] for k in range(4)]
    return r


def rotation(q):
    x, y, z, w = q
    qxx, qyy, qzz = x * x, y * y, z * z
    qxz, qxy, qyz = x * z, x * y, y * z
    qwx, qwy, qwz = w * x, w * y, w * z
    one, two = F(1), F(2)
    r = identity()
    r[0][0] = one - two * (qyy + qzz)
    r[0][1] = two * (qxy + qwz)
    r[0][2] = two * (qxz - qwy)
    r[1][0] = two * (qxy - qwz)
    r[1][1] = one - two * (qxx + qzz)
    r[1][2] = two * (qyz + qwx)
    r[2][0] = two * (qxz + qwy)
    r[2][1] = two * (qyz - qwx)
    r[2][2] = one - two * (qxx + qyy)
    return r


def inverse(m):
    c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3]
    c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3]
    c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3]
    c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3]
    c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3]
    c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3]
    c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2]
    c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2]
    c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2]
    c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3]
    c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3]
    c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3]
    c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2]
    c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2]
    c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2]
    c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1]
    c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1]
    c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1]
    fac = [[c00, c00, c02, c03], [c04, c04, c06, c07], [c08, c08, c10, c11],
           [c12, c12, c14, c15], [c16, c16, c18, c19], [c20, c20, c22, c23]]
    vec = [[m[1][r], m[0][r], m[0][r], m[0][r]] for r in range(4)]
    v0, v1, v2, v3 = vec
    f0, f1, f2, f3, f4, f5 = fac
    inv0 = [(v1[k] * f0[k] - v2[k] * f1[k]) + v3[k] * f2[k] for k in range(4)]
    inv1 = [(v0[k] * f0[k] - v2[k] * f3[k]) + v3[k] * f4[k] for k in range(4)]
    inv2 = [(v0[k] * f1[k] - v1[k] * f3[k]) + v3[k] * f5[k] for k in range(4)]
    inv3 = [(v0[k] * f2[k] - v1[k] * f4[k]) + v2[k] * f5[k] for k in range(4)]
    sign_a = [F(1), F(-1), F(1), F(-1)]
    sign_b = [F(-1), F(1), F(-1), F(1)]
    inv = [[inv0[k] * sign_a[k] for k in range(4)], [inv1[k] * sign_b[k] for k in range(4)],
           [inv2[k] * sign_a[k] for k in range(4)], [inv3[k] * sign_b[k] for k in range(4)]]
    row0 = [inv[0][0], inv[1][0], inv[2][0], inv[3][0]]
    dot0 = [m[0][k] * row0[k] for k in range(4)]
    det = (dot0[0] + dot0[1]) + (dot0[2] + dot0[3])
    s = F(1) / det
    return [[inv[c][k] * s for k in range(4)] for c in range(4)]


def transpose(m):
    return [[m[k][c] for k in range(4)] for c in range(4)]


def normalize3(v):
    with np.errstate(divide="ignore", invalid="ignore"):
        d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
        s = F(1) / np.sqrt(d)
        return [v[0] * s, v[1] * s, v[2] * s]


def _buffers(doc, base_dir):
    out = []
    for b in doc.get("buffers", []):
        uri = b["uri"]
        if uri.startswith("data:"):
            data = base64.b64decode(uri.split(",", 1)[1])
        else:
            with open(os.path.join(base_dir, urllib.parse.unquote(uri)), "rb") as f:
                data = f.read()
        out.append(data)
    return out


def _data(doc, bufs, acc, nbytes):
    view = doc["bufferViews"][acc["bufferView"]]
    off = acc.get("byteOffset", 0) + view.get("byteOffset", 0)
    return bufs[view["buffer"]][off:off + nbytes]


def load_model(path, untextured=False):
    """(vertices, indices, materials) of one file, model-local indices."""
    with open(path, "rb") as f:
        doc = json.loads(f.read())
    bufs = _buffers(doc, os.path.dirname(path))
    mats = np.zeros(len(doc.get("materials", [])), MATERIAL_DTYPE)
    for i, m in enumerate(doc.get("materials", [])):
        pbr = m.get("pbrMetallicRoughness", {})
        mats[i]["base_color_texture_index"] = -1
        if "baseColorFactor" in pbr:
            c = pbr["baseColorFactor"]
            mats[i]["base_color"] = [F(c[0]), F(c[1]), F(c[2]), F(c[3]) if len(c) > 3 else F(1)]
        if "baseColorTexture" in pbr:
            mats[i]["base_color_texture_index"] = pbr["baseColorTexture"].get("index", -1)
        if untextured:
            mats[i]["base_color_texture_index"] = -1
        mats[i]["material_model"] = DIFFUSE_LIGHT if m.get("name") == "light" else LAMBERTIAN
    verts, idx = [], []

    def node_matrix(n):
        m = identity()
        if len(n.get("translation", [])) == 3:
            m = translate(m, [F(x) for x in n["translation"]])
        if len(n.get("rotation", [])) == 4:
            m = mul_mm(m, rotation([F(x) for x in n["rotation"]]))
        if len(n.get("scale", [])) == 3:
            m = scale(m, [F(x) for x in n["scale"]])
        if len(n.get("matrix", [])) == 16:
            a = [F(x) for x in n["matrix"]]
            m = [a[4 * c:4 * c + 4] for c in range(4)]
        return m

    def visit(i, chain):
        n = doc["nodes"][i]
        mine = node_matrix(n)
        for c in n.get("children", []):
            visit(c, [mine] + chain)
        if "mesh" not in n:
            return
        world = mine
        for p in chain:  # nearest parent first
            world = mul_mm(p, world)
        nrm = transpose(inverse(world))
        for prim in doc["meshes"][n["mesh"]]["primitives"]:
            start = len(verts)
            a = doc["accessors"][prim["attributes"]["POSITION"]]
            cnt = a["count"]
            pos = np.frombuffer(_data(doc, bufs, a, cnt * 12), "<f4").reshape(cnt, 3)
            nor = uv = None
            if "NORMAL" in prim["attributes"]:
                a = doc["accessors"][prim["attributes"]["NORMAL"]]
                nor = np.frombuffer(_data(doc, bufs, a, cnt * 12), "<f4").reshape(cnt, 3)
            if "TEXCOORD_0" in prim["attributes"]:
                a = doc["accessors"][prim["attributes"]["TEXCOORD_0"]]
                uv = np.frombuffer(_data(doc, bufs, a, cnt * 8), "<f4").reshape(cnt, 2)
            mat = prim.get("material", -1)
            for v in range(cnt):
                p = mul_mv(world, [pos[v, 0], pos[v, 1], pos[v, 2], F(1)])
                n3 = [nor[v, 0], nor[v, 1], nor[v, 2]] if nor is not None else [F(0), F(0), F(0)]
                nn = normalize3(n3)
                q = mul_mv(nrm, [nn[0], nn[1], nn[2], F(1)])
                t = (uv[v, 0], uv[v, 1]) if uv is not None else (F(0), F(0))
                verts.append(((p[0], p[1], p[2]), (q[0], q[1], q[2]), t, mat))
            a = doc["accessors"][prim["indices"]]
            size = {5125: 4, 5123: 2, 5121: 1}[a["componentType"]]
            raw = np.frombuffer(_data(doc, bufs, a, a["count"] * size), {4: "<u4", 2: "<u2", 1: "u1"}[size])
            idx.extend(int(x) + start for x in raw)

    for r in doc["scenes"][0]["nodes"]:
        visit(r, [])
    v = np.array(verts, VERTEX_DTYPE) if verts else np.zeros(0, VERTEX_DTYPE)
    return v, np.array(idx, np.uint32), mats


def load(paths, untextured=False):
    """AScene arrays (vertices, indices, materials, offsets) of several files."""
    vs, is_, ms, offs = [], [], [], []
    nv = ni = nm = 0
    for p in paths:
        v, i, m = load_model(p, untextured)
        offs += [ni, nv]
        v = v.copy()
        v["material_index"] += nm
        vs.append(v)
        is_.append(i)
        ms.append(m)
        nv, ni, nm = nv + len(v), ni + len(i), nm + len(m)
    return (np.concatenate(vs), np.concatenate(is_), np.concatenate(ms), np.array(offs, np.uint32))
