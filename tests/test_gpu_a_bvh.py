"""Structure of the GPU-built BVH, checked on the host before any traversal
test runs (this file is collected first).  SURVEY-style validity criteria:
every triangle sits in exactly one leaf, every node is reached exactly once,
and each stored child box contains everything below it."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import cornell_scene, say, small_terrain
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def check_bvh(nodes: np.ndarray, tris: np.ndarray, tri_count: int) -> int:
    links = nodes[:, 12:14].copy().view(np.int32)
    n_nodes = len(nodes)
    assert len(tris) == tri_count
    tv = tris.reshape(-1, 3, 4)[:, :, :3]
    tlo, thi = tv.min(axis=1), tv.max(axis=1)
    covered = np.zeros(tri_count, np.int64)
    seen = np.zeros(n_nodes, np.int64)
    seen[0] = 1
    boxes = {}

    def child_box(i, side):
        a, b, c = nodes[i, 0:4], nodes[i, 4:8], nodes[i, 8:12]
        if side == 0:
            return np.array([a[0], a[2], c[0]]), np.array([a[1], a[3], c[1]])
        return np.array([b[0], b[2], c[2]]), np.array([b[1], b[3], c[3]])

    # iterative DFS with depth
    stack = [(0, 0)]
    max_depth = 0
    while stack:
        i, d = stack.pop()
        for side in range(2):
            link = int(links[i, side])
            lo, hi = child_box(i, side)
            if link >= 0:
                assert link < n_nodes, f"node {i} links to {link}"
                seen[link] += 1
                assert seen[link] == 1 or tri_count == 1, f"node {link} reached twice"
                # child's own children boxes must be inside this box
                for s2 in range(2):
                    l2, h2 = child_box(link, s2)
                    assert np.all(l2 >= lo) and np.all(h2 <= hi), f"box of node {link} not contained"
                stack.append((link, d + 1))
            else:
                code = ~link
                first, cnt = code >> 4, (code & 15) + 1
                assert 0 <= first and first + cnt <= tri_count
                covered[first:first + cnt] += 1
                assert np.all(tlo[first:first + cnt] >= lo) and np.all(thi[first:first + cnt] <= hi)
                max_depth = max(max_depth, d + 1)
    # collapsed subtrees leave their internal nodes unlinked; no node is shared
    assert (seen <= 1).all() or tri_count == 1, "shared nodes"
    if tri_count == 1:
        assert covered[0] == 2
    else:
        assert (covered == 1).all(), "triangle not in exactly one leaf"
    ids = tris[:, 3].copy().view(np.uint32)
    assert np.array_equal(np.sort(ids), np.arange(tri_count, dtype=np.uint32)), "canonical ids not a permutation"
    return max_depth


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH])
@pytest.mark.parametrize("which", ["cornell", "terrain"])
def test_bvh_structure(which, builder):
    scene = cornell_scene() if which == "cornell" else small_terrain()
    say(f"[bvh] building {which} ({scene.triangle_count} tris) with builder {builder}")
    ctx = N.Context(scene, 8, 8, builder=builder)
    info = ctx.accel_info()
    assert info.builder == builder
    nodes, tris = ctx.read_accel()
    depth = check_bvh(nodes, tris, scene.triangle_count)
    # the builder reports the uncollapsed depth, an upper bound on the linked tree's
    assert 0 < depth <= info.max_depth
    # canonical triangle vertices match the scene's index buffer
    v, idx, _, off = scene.arrays()
    ids = tris[:, 3].copy().view(np.uint32)
    order = np.argsort(ids)
    canon = tris[order].reshape(-1, 3, 4)[:, :, :3]
    inst = tris[order, 7].copy().view(np.uint32)
    prim = tris[order, 11].copy().view(np.uint32)
    vi = off[2 * inst + 1][:, None] + idx[(off[2 * inst] + 3 * prim)[:, None] + np.arange(3)]
    assert np.array_equal(canon, v["pos"][vi])


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH])
def test_bvh_two_triangles_and_rebuild(builder):
    """Smallest clustered tree, and a rebuild of one context with the other builder."""
    verts = np.zeros(6, N.VERTEX_DTYPE)
    verts["pos"] = [[0, 0, 1], [1, 0, 1], [0, 1, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2]]
    mats = np.zeros(1, N.MATERIAL_DTYPE)
    mats["base_color_texture_index"] = -1
    scene = N.Scene.from_arrays(verts, np.arange(6, dtype=np.uint32), mats, np.array([0, 0], np.uint32))
    ctx = N.Context(scene, 4, 4, builder=builder)
    check_bvh(*ctx.read_accel(), 2)
    ctx.build_accel(N.XR_BUILDER_PLOC if builder == N.XR_BUILDER_LBVH else N.XR_BUILDER_LBVH)
    check_bvh(*ctx.read_accel(), 2)
    r = np.zeros(1, N.RAY_DTYPE)
    r["origin"] = [[0.2, 0.2, 0.0]]
    r["direction"] = [0, 0, 1]
    r["tmin"], r["tmax"] = 1e-4, 1e4
    h = ctx.trace(r)
    assert h["t"][0] == np.float32(1.0) and h["triangle"][0] == 0


def test_bvh_single_triangle():
    verts = np.zeros(3, N.VERTEX_DTYPE)
    verts["pos"] = [[0, 0, 1], [1, 0, 1], [0, 1, 1]]
    verts["normal"] = [0, 0, -1]
    mats = np.zeros(1, N.MATERIAL_DTYPE)
    mats["base_color"] = [0.5, 0.5, 0.5, 1]
    mats["base_color_texture_index"] = -1
    scene = N.Scene.from_arrays(verts, np.arange(3, dtype=np.uint32), mats, np.array([0, 0], np.uint32))
    ctx = N.Context(scene, 4, 4)
    nodes, tris = ctx.read_accel()
    check_bvh(nodes, tris, 1)
    r = np.zeros(2, N.RAY_DTYPE)
    r["origin"] = [[0.2, 0.2, 0.0], [0.9, 0.9, 0.0]]
    r["direction"] = [0, 0, 1]
    r["tmin"], r["tmax"] = 1e-4, 1e4
    h = ctx.trace(r)
    assert h["t"][0] == np.float32(1.0) and h["triangle"][0] == 0
    assert h["t"][1] == -1.0 and h["triangle"][1] == 0xFFFFFFFF
