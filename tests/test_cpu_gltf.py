"""glTF ingestion (xrh_scene_gltf, no GPU needed).

The loader must give the AScene arrays the reference's AglTFModel / AScene
build (src/Asset/Model.cpp:69-260, src/Asset/Scene.cpp:16-50).  Checked
three ways:
  * bit for bit against oracle/gltf.py, the float32 restatement of that code
    and of glm's operation order;
  * against closed forms in float64 (rotations, TRS nodes, parent chains,
    inverse-transpose normals), which a dropped term, a transposed matrix or
    a wrong multiplication order would fail;
  * against what the reference's own cb2 asset records (accessor min/max,
    primitive counts, the emitter material).
"""
from __future__ import annotations

import base64
import json
import math
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import gltf as G
from xvk_ray_amd import native as N

CB2 = Path(__file__).parent / "golden" / "gltf" / "cb2" / "scene.gltf"
LAMBERTIAN, DIFFUSE_LIGHT = 0, 4  # AMaterial::Enum (src/Asset/Material.h:36-42)


def _same_bytes(a: np.ndarray, b: np.ndarray) -> bool:
    return a.tobytes() == b.tobytes()


def _check_against_restatement(paths, untextured=False):
    scene = N.Scene.gltf(*paths, untextured=untextured)
    v, i, m, o = scene.arrays()
    rv, ri, rm, ro = G.load([str(p) for p in paths], untextured)
    assert len(v) == len(rv) and _same_bytes(v, rv.astype(v.dtype)), "vertices"
    assert np.array_equal(i, ri), "indices"
    assert _same_bytes(m, rm.astype(m.dtype)), "materials"
    assert np.array_equal(o, ro), "offsets"
    return v, i, m, o


# ------------------------------------------------------------------ cb2 ----
def test_cb2_matches_restatement():
    _check_against_restatement([CB2])


def test_cb2_facts_from_the_file():
    v, i, m, o = N.Scene.gltf(CB2).arrays()
    doc = json.loads(CB2.read_text())
    prims = [p for mesh in doc["meshes"] for p in mesh["primitives"]]
    assert len(i) == sum(doc["accessors"][p["indices"]]["count"] for p in prims) == 96
    assert len(v) == sum(doc["accessors"][p["attributes"]["POSITION"]]["count"] for p in prims) == 64
    assert list(o) == [0, 0]
    names = [mat["name"] for mat in doc["materials"]]
    for k, name in enumerate(names):
        want = DIFFUSE_LIGHT if name == "light" else LAMBERTIAN
        assert m[k]["material_model"] == want
        assert np.array_equal(m[k]["base_color"],
                              np.array(doc["materials"][k]["pbrMetallicRoughness"]["baseColorFactor"], np.float32))
        assert m[k]["base_color_texture_index"] == -1
    # the scene's root node is a +90 degree turn about x: local (x, y, z) -> (x, z, -y).
    # The world box is the accessors' recorded min/max taken through that turn.
    lo = np.min([doc["accessors"][p["attributes"]["POSITION"]]["min"] for p in prims], axis=0)
    hi = np.max([doc["accessors"][p["attributes"]["POSITION"]]["max"] for p in prims], axis=0)
    pos = v["pos"].astype(np.float64)
    want_lo = np.array([lo[0], lo[2], -hi[1]])
    want_hi = np.array([hi[0], hi[2], -lo[1]])
    assert np.allclose(pos.min(0), want_lo, atol=1e-6) and np.allclose(pos.max(0), want_hi, atol=1e-6)
    # every vertex of a primitive carries that primitive's material, in node order
    mats = []
    for node in doc["nodes"][2:]:
        p = doc["meshes"][node["mesh"]]["primitives"][0]
        mats += [p["material"]] * doc["accessors"][p["attributes"]["POSITION"]]["count"]
    assert list(v["material_index"]) == mats


def test_cb2_closed_form_turn():
    """World positions and normals are the local ones turned by the root
    matrix (x, y, z) -> (x, z, -y); normals stay unit length."""
    doc = json.loads(CB2.read_text())
    raw = (CB2.parent / "scene.bin").read_bytes()
    v, _, _, _ = N.Scene.gltf(CB2).arrays()
    k = 0
    for node in doc["nodes"][2:]:
        p = doc["meshes"][node["mesh"]]["primitives"][0]
        for attr, field in (("POSITION", "pos"), ("NORMAL", "normal")):
            a = doc["accessors"][p["attributes"][attr]]
            view = doc["bufferViews"][a["bufferView"]]
            off = view.get("byteOffset", 0) + a.get("byteOffset", 0)
            loc = np.frombuffer(raw[off:off + 12 * a["count"]], "<f4").reshape(-1, 3).astype(np.float64)
            if attr == "NORMAL":
                loc = loc / np.linalg.norm(loc, axis=1, keepdims=True)
            want = np.stack([loc[:, 0], loc[:, 2], -loc[:, 1]], axis=1)
            got = v[field][k:k + a["count"]].astype(np.float64)
            assert np.allclose(got, want, atol=2e-7), (attr, got[:2], want[:2])
        k += a["count"]
    assert np.allclose(np.linalg.norm(v["normal"].astype(np.float64), axis=1), 1.0, atol=1e-6)


# ------------------------------------------------------------ synthetic ----
def _write(tmp: Path, name: str, doc: dict, blobs: list[bytes], data_uri: bool = False) -> Path:
    doc = dict(doc)
    doc["asset"] = {"version": "2.0"}
    doc["buffers"] = []
    for k, b in enumerate(blobs):
        if data_uri:
            uri = "data:application/octet-stream;base64," + base64.b64encode(b).decode()
        else:
            fname = f"{name} buffer {k}.bin"  # a space: the uri is percent-encoded
            (tmp / fname).write_bytes(b)
            uri = fname.replace(" ", "%20")
        doc["buffers"].append({"uri": uri, "byteLength": len(b)})
    path = tmp / f"{name}.gltf"
    path.write_text(json.dumps(doc))
    return path


class _Blob:
    """Builds one buffer plus views/accessors."""

    def __init__(self):
        self.data = bytearray()
        self.views, self.accessors = [], []

    def add(self, arr: np.ndarray, ctype: int, typ: str) -> int:
        while len(self.data) % 4:
            self.data.append(0)
        self.views.append({"buffer": 0, "byteOffset": len(self.data), "byteLength": arr.nbytes})
        self.data += arr.tobytes()
        self.accessors.append({"bufferView": len(self.views) - 1, "componentType": ctype,
                               "count": int(arr.shape[0]), "type": typ})
        return len(self.accessors) - 1


def _quad(blob: _Blob, index_type: int, normals=True, uvs=False, material=None, z=0.0):
    pos = np.array([[0, 0, z], [1, 0, z], [1, 1, z], [0, 1, z]], np.float32)
    attrs = {"POSITION": blob.add(pos, 5126, "VEC3")}
    if normals:  # not unit length: the loader normalises them
        nrm = np.array([[0, 0, 2], [0.3, -0.7, 1.9], [-1.1, 0.2, 0.9], [0.05, 0.6, 1.3]], np.float32)
        attrs["NORMAL"] = blob.add(nrm, 5126, "VEC3")
    if uvs:
        attrs["TEXCOORD_0"] = blob.add(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), 5126, "VEC2")
    dt = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[index_type]
    prim = {"attributes": attrs, "indices": blob.add(np.array([0, 1, 2, 0, 2, 3], dt), index_type, "SCALAR")}
    if material is not None:
        prim["material"] = material
    return prim


def _axis_angle(axis, deg):
    h = math.radians(deg) / 2
    a = np.array(axis, float) / np.linalg.norm(axis)
    return [*(a * math.sin(h)), math.cos(h)]


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


QUAD_NORMALS = np.array([[0, 0, 2], [0.3, -0.7, 1.9], [-1.1, 0.2, 0.9], [0.05, 0.6, 1.3]], np.float32)
# a general node matrix, column major: shear, non-uniform scale, translation
# and a non-zero bottom row (glTF asks for affine matrices; the loader, like
# glm, handles any), so every term of the 4x4 inverse matters
SHEAR = [1.3, 0.2, -0.4, 0.1, 0.7, 0.9, 0.35, -0.05, -0.25, 0.6, 1.7, 0.02, 0.5, -1.25, 2.0, 1]


def _synthetic(tmp: Path, data_uri=False) -> tuple[Path, dict]:
    b = _Blob()
    meshes = [
        {"primitives": [_quad(b, 5125, material=0), _quad(b, 5123, uvs=True, material=1, z=0.5)]},
        {"primitives": [_quad(b, 5121, material=2)]},
        {"primitives": [_quad(b, 5125, normals=False)]},
        {"primitives": [_quad(b, 5125, material=2)]},
    ]
    q = _axis_angle([0.3, 1.0, -0.2], 37.0)
    doc = {
        "scene": 0,
        "scenes": [{"nodes": [0, 3, 4]}],
        "nodes": [
            {"translation": [1.5, -2.0, 0.25], "rotation": q, "scale": [2.0, 0.5, 3.0], "children": [1, 2],
             "mesh": 0},
            {"translation": [0.0, 1.0, 0.0], "mesh": 1},
            {"matrix": [0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 4, 5, 6, 1], "translation": [9, 9, 9]},
            {"mesh": 2},
            {"matrix": SHEAR, "mesh": 3},
        ],
        "meshes": meshes,
        "materials": [
            {"name": "light", "pbrMetallicRoughness": {"baseColorFactor": [3.75, 3.75, 3.75, 1.0]}},
            {"name": "rgb", "pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.2, 0.3],
                                                      "baseColorTexture": {"index": 2}}},
            {"name": "plain"},
        ],
        "bufferViews": b.views,
        "accessors": b.accessors,
    }
    return _write(tmp, "synthetic", doc, [bytes(b.data)], data_uri), {"q": q}


@pytest.mark.parametrize("data_uri", [False, True])
def test_synthetic_matches_restatement(tmp_path, data_uri):
    path, _ = _synthetic(tmp_path, data_uri)
    _check_against_restatement([path])


def test_synthetic_closed_forms(tmp_path):
    path, info = _synthetic(tmp_path)
    v, i, m, o = N.Scene.gltf(path).arrays()
    pos = v["pos"].astype(np.float64)
    nrm = v["normal"].astype(np.float64)
    uv = v["tex_coord"]
    mat = v["material_index"]
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    R = _rot(info["q"])
    S = np.diag([2.0, 0.5, 3.0])
    T = np.array([1.5, -2.0, 0.25])
    root = lambda p: (R @ (S @ p.T)).T + T  # noqa: E731
    # children first: node 1 (child of 0), node 2 (no mesh), then node 0's own
    # mesh, then root node 3
    child = root(quad + [0.0, 1.0, 0.0])
    own0, own1 = root(quad), root(quad + [0, 0, 0.5])
    M = np.array(SHEAR, float).reshape(4, 4).T  # column major -> row major
    sheared = (M @ np.c_[quad, np.ones(4)].T).T[:, :3]  # xyz of M (p, 1), no divide
    want = np.concatenate([child, own0, own1, quad, sheared])
    assert np.allclose(pos, want, atol=1e-5), np.abs(pos - want).max()
    # normals: inverse-transpose of the linear part applied to the unit normal, not renormalised
    unit = QUAD_NORMALS.astype(float) / np.linalg.norm(QUAD_NORMALS.astype(float), axis=1, keepdims=True)
    n_root = (np.linalg.inv(R @ S).T @ unit.T).T
    assert np.allclose(nrm[:12], np.tile(n_root, (3, 1)), atol=1e-5)
    assert np.isnan(nrm[12:16]).all()  # no NORMAL attribute: normalize(0) = NaN, as in the reference
    assert np.allclose(nrm[16:], (np.linalg.inv(M).T @ np.c_[unit, np.ones(4)].T).T[:, :3], atol=1e-5)
    # texture coordinates only on node 0's second primitive
    assert np.array_equal(uv[8:12], np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    assert not uv[:8].any() and not uv[12:].any()
    assert list(mat) == [2] * 4 + [0] * 4 + [1] * 4 + [-1] * 4 + [2] * 4
    # indices: primitive-local plus the primitive's first vertex, u8/u16/u32 alike
    assert list(i) == [k + s for s in (0, 4, 8, 12, 16) for k in (0, 1, 2, 0, 2, 3)]
    # materials: factor, 3-component factor gets alpha 1, missing factor is zero,
    # "light" is the emitter, texture index kept
    assert m[0]["material_model"] == DIFFUSE_LIGHT and m[1]["material_model"] == m[2]["material_model"] == LAMBERTIAN
    assert np.array_equal(m[1]["base_color"], np.array([0.1, 0.2, 0.3, 1.0], np.float32))
    assert not m[2]["base_color"].any()
    assert [int(x) for x in m["base_color_texture_index"]] == [-1, 2, -1]
    assert list(o) == [0, 0]


def test_node_matrix_wins_over_trs(tmp_path):
    """A node with both `matrix` and TRS uses the matrix (Model.cpp:116-118)."""
    b = _Blob()
    doc = {"scenes": [{"nodes": [0]}],
           "nodes": [{"matrix": [0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 4, 5, 6, 1], "translation": [9, 9, 9],
                      "mesh": 0}],
           "meshes": [{"primitives": [_quad(b, 5125)]}], "bufferViews": b.views, "accessors": b.accessors}
    path = _write(tmp_path, "matrix", doc, [bytes(b.data)])
    v, _, _, _ = N.Scene.gltf(path).arrays()
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    want = np.stack([-quad[:, 1] + 4, quad[:, 0] + 5, quad[:, 2] + 6], axis=1)
    assert np.array_equal(v["pos"].astype(np.float64), want)
    _check_against_restatement([path])


def test_several_files_offsets_and_materials(tmp_path):
    a, _ = _synthetic(tmp_path)
    v, i, m, o = _check_against_restatement([a, CB2])
    nv_a, ni_a, nm_a = 20, 30, 3
    assert list(o) == [0, 0, ni_a, nv_a]
    mats = v["material_index"]
    assert list(mats[nv_a:nv_a + 4]) == [0 + nm_a] * 4
    assert mats[12] == -1 and len(m) == nm_a + 8  # -1 stays -1 in the first file (offset 0)


def test_untextured_flag(tmp_path):
    path, _ = _synthetic(tmp_path)
    _, _, m, _ = _check_against_restatement([path], untextured=True)
    assert (m["base_color_texture_index"] == -1).all()


def test_texture_index_without_textures_is_kept(tmp_path):
    """Material 1 names texture 2 but the file has no textures: the index is
    kept as the reference keeps it (a frame then fails with XR_ERR_STATE)."""
    path, _ = _synthetic(tmp_path)
    scene = N.Scene.gltf(path)
    assert scene.textures() == []
    assert scene.arrays()[2]["base_color_texture_index"][1] == 2


@pytest.mark.parametrize("case", ["missing", "json", "float_indices", "no_indices", "short_buffer", "past_buffer",
                                  "count_wraps", "count_huge", "index_count_wraps", "count_fraction", "material_huge"])
def test_errors(tmp_path, case):
    if case == "missing":
        path = tmp_path / "nope.gltf"
    elif case == "json":
        path = tmp_path / "bad.gltf"
        path.write_text('{"scenes": [ {"nodes": [0] ]}')
    else:
        b = _Blob()
        prim = _quad(b, 5125)
        if case == "float_indices":
            b.accessors[prim["indices"]]["componentType"] = 5126
        if case == "no_indices":
            del prim["indices"]
        if case == "past_buffer":
            b.accessors[prim["attributes"]["POSITION"]]["count"] = 1000
        if case == "count_wraps":  # count * 12 wraps to 0 in 64 bits
            b.accessors[prim["attributes"]["POSITION"]]["count"] = 2 ** 62
        if case == "count_huge":
            b.accessors[prim["attributes"]["POSITION"]]["count"] = 1e300
        if case == "index_count_wraps":  # count * 4 wraps to 0 in 64 bits
            b.accessors[prim["indices"]]["count"] = 2 ** 62
        if case == "count_fraction":
            b.accessors[prim["attributes"]["POSITION"]]["count"] = 3.5
        if case == "material_huge":
            prim["material"] = 1e12
        doc = {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}], "meshes": [{"primitives": [prim]}],
               "bufferViews": b.views, "accessors": b.accessors}
        path = _write(tmp_path, case, doc, [bytes(b.data)])
        if case == "short_buffer":
            d = json.loads(path.read_text())
            d["buffers"][0]["byteLength"] += 100
            path.write_text(json.dumps(d))
    with pytest.raises(N.XrError) as e:
        N.Scene.gltf(path)
    assert e.value.code == 1 and "xrh_scene_gltf:" in str(e.value)


def test_struct_layout_matches_reference():
    """AVertex 36 B and AMaterial 32 B, as the arrays the loader fills."""
    v, _, m, _ = N.Scene.gltf(CB2).arrays()
    assert v.dtype.itemsize == 36 and m.dtype.itemsize == 32
    assert struct.calcsize("<3f3f2fi") == 36


def test_loader_survives_damaged_files(tmp_path):
    """cb2's glTF with JSON fields replaced by out-of-range, wrong-typed or
    huge values (counts, offsets, component types, indices), fields deleted,
    text damaged, and its buffer truncated: every load returns a scene or an
    error, never crashes."""
    import copy
    import random
    import shutil
    src = Path(__file__).parent / "golden" / "gltf" / "cb2"
    base = json.loads((src / "scene.gltf").read_text())
    blob = (src / "scene.bin").read_bytes()
    vals = [0, 1, -1, 3, 4, 5, 12, 255, 256, 65535, 2 ** 31 - 1, 2 ** 31, 2 ** 32, 2 ** 62, 1e300, -1e300, 1.5,
            5121, 5123, 5125, 5126, "VEC3", "SCALAR", None, [], {}, "x", True]

    def paths(o, pre=()):
        if isinstance(o, dict):
            for k, v in o.items():
                yield pre + (k,)
                yield from paths(v, pre + (k,))
        elif isinstance(o, list):
            for i, v in enumerate(o):
                yield pre + (i,)
                yield from paths(v, pre + (i,))

    every = list(paths(base))
    rng = random.Random(41)
    loaded = 0
    for it in range(400):
        g = copy.deepcopy(base)
        for _ in range(rng.randint(1, 3)):
            p = rng.choice(every)
            try:
                o = g
                for k in p[:-1]:
                    o = o[k]
                if rng.random() < 0.15:
                    del o[p[-1]]
                else:
                    o[p[-1]] = rng.choice(vals)
            except (KeyError, IndexError, TypeError):
                pass
        text = json.dumps(g)
        if rng.random() < 0.1:
            t = list(text)
            for _ in range(rng.randint(1, 4)):
                t[rng.randrange(len(t))] = rng.choice('{}[]",:0123456789.e- ')
            text = "".join(t)
        d = tmp_path / f"s{it % 4}"
        d.mkdir(exist_ok=True)
        (d / "scene.gltf").write_text(text)
        (d / "scene.bin").write_bytes(blob if rng.random() > 0.1 else blob[:rng.randrange(len(blob))])
        try:
            N.Scene.gltf(d / "scene.gltf")
            loaded += 1
        except N.XrError:
            pass
    shutil.rmtree(tmp_path, ignore_errors=True)
    assert 0 < loaded < 400
