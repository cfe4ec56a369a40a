"""Structure of the GPU-built BVH, checked on the host before any traversal
test runs (this file is collected first).  Validity criteria:
every triangle sits in exactly one leaf, every node is reached exactly once,
and each stored child box contains everything below it."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import cornell_scene, say, small_terrain
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def decode_nodes(nodes: np.ndarray):
    """(lo, hi, links, counts) of the 64-byte four-child nodes: lo / hi as
    (n, 4, 3) float64 boxes of the children, decoded as origin + g * step."""
    origin = nodes[:, 0:3].astype(np.float64)
    step = np.stack([nodes[:, 3], nodes[:, 10], nodes[:, 11]], axis=1).astype(np.float64)  # (n, 3)
    mant, _ = np.frexp(step)
    assert (mant == 0.5).all(), "grid steps are powers of two"
    planes = nodes[:, 4:10].copy().view(np.uint32)  # lox, hix, loy, hiy, loz, hiz
    g = np.stack([(planes >> (8 * k)) & 0xFF for k in range(4)], axis=1).astype(np.int64)  # (n, 4, 6)
    links = nodes[:, 12:16].copy().view(np.int32)
    # used slots come first; an unused slot is the empty box 255 / 0 with link -1
    empty = (g[:, :, 0::2] == 255).all(axis=2) & (g[:, :, 1::2] == 0).all(axis=2)
    counts = (~empty).sum(axis=1)
    assert (empty == (np.arange(4)[None, :] >= counts[:, None])).all(), "unused slots follow the children"
    assert (links[empty] == -1).all(), "unused slots link the leaf of triangle 0"
    lo = origin[:, None, :] + g[:, :, 0::2] * step[:, None, :]
    hi = origin[:, None, :] + g[:, :, 1::2] * step[:, None, :]
    return lo, hi, links, counts


def check_bvh(nodes: np.ndarray, tris: np.ndarray, tri_count: int) -> int:
    """Walks the four-child tree from the root: every stored child box holds
    every triangle below it, every triangle sits in exactly one leaf, every
    node is reached once.  Returns the number of nodes on the longest
    root-to-leaf path."""
    assert nodes.shape[1] == 16
    lo, hi, links, counts = decode_nodes(nodes)
    n_nodes = len(nodes)
    assert len(tris) == tri_count
    assert ((counts >= 2) & (counts <= 4)).all(), "child counts"
    tv = tris.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    tlo, thi = tv.min(axis=1), tv.max(axis=1)
    covered = np.zeros(tri_count, np.int64)
    seen = np.zeros(n_nodes, np.int64)
    seen[0] = 1
    # subtree triangle bounds, children before parents (levels are stored in order)
    sub_lo = np.full((n_nodes, 3), np.inf)
    sub_hi = np.full((n_nodes, 3), -np.inf)
    depth = np.zeros(n_nodes, np.int64)
    depth[0] = 1
    max_depth = 0
    order = []
    stack = [0]
    while stack:
        i = stack.pop()
        order.append(i)
        for k in range(int(counts[i])):
            link = int(links[i, k])
            if link >= 0:
                assert link < n_nodes, f"node {i} links to {link}"
                seen[link] += 1
                assert seen[link] == 1, f"node {link} reached twice"
                depth[link] = depth[i] + 1
                stack.append(link)
            else:
                code = ~link
                first, n = code >> 4, (code & 15) + 1
                assert 0 <= first and first + n <= tri_count
                covered[first:first + n] += 1
                assert np.all(tlo[first:first + n] >= lo[i, k]) and np.all(thi[first:first + n] <= hi[i, k]), \
                    f"leaf box {k} of node {i} does not hold its triangles"
                sub_lo[i] = np.minimum(sub_lo[i], tlo[first:first + n].min(axis=0))
                sub_hi[i] = np.maximum(sub_hi[i], thi[first:first + n].max(axis=0))
                max_depth = max(max_depth, int(depth[i]))
    for i in reversed(order):  # children are finished before their parent
        for k in range(int(counts[i])):
            link = int(links[i, k])
            if link >= 0:
                assert np.all(sub_lo[link] >= lo[i, k]) and np.all(sub_hi[link] <= hi[i, k]), \
                    f"box {k} of node {i} does not hold the triangles below it"
                sub_lo[i] = np.minimum(sub_lo[i], sub_lo[link])
                sub_hi[i] = np.maximum(sub_hi[i], sub_hi[link])
    assert (seen == 1).all(), "nodes not reached exactly once"
    if tri_count == 1:
        assert covered[0] == 2
    else:
        assert (covered == 1).all(), "triangle not in exactly one leaf"
    ids = tris[:, 3].copy().view(np.uint32)
    assert np.array_equal(np.sort(ids), np.arange(tri_count, dtype=np.uint32)), "canonical ids not a permutation"
    return max_depth


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH, N.XR_BUILDER_SAH_HOST])
@pytest.mark.parametrize("which", ["cornell", "terrain"])
def test_bvh_structure(which, builder):
    scene = cornell_scene() if which == "cornell" else small_terrain()
    say(f"[bvh] building {which} ({scene.triangle_count} tris) with builder {builder}")
    ctx = N.Context(scene, 8, 8, builder=builder)
    info = ctx.accel_info()
    assert info.builder == builder
    nodes, tris = ctx.read_accel()
    depth = check_bvh(nodes, tris, scene.triangle_count)
    assert depth == info.max_depth
    assert info.node_bytes == 64 * len(nodes)
    # the collapse opens inner children until a node has four: a node with
    # fewer children has leaves only
    lo, hi, links, counts = decode_nodes(nodes)
    short = counts < 4
    assert short.mean() < 0.9
    assert (links[short][:, :2] < 0).all() and ((links[short][:, 2] < 0) | (counts[short] < 3)).all()
    # the grid spans the children: per axis the lowest plane is 0 (the origin
    # is the children's min corner) and, the step being the smallest power of
    # two that reaches the max corner, the highest plane lies in the grid's
    # upper half
    planes = nodes[:, 4:10].copy().view(np.uint32)
    g = np.stack([(planes >> (8 * k)) & 0xFF for k in range(4)], axis=1)  # (n, 4, 6)
    live = np.arange(4)[None, :] < counts[:, None]
    step = np.stack([nodes[:, 3], nodes[:, 10], nodes[:, 11]], axis=1).astype(np.float64)
    for a in range(3):
        g_lo = np.where(live, g[:, :, 2 * a], 255).min(axis=1)
        g_hi = np.where(live, g[:, :, 2 * a + 1], 0).max(axis=1)
        assert (g_lo == 0).all()
        wide = step[:, a] > 2.0**-126  # the smallest step serves children of zero extent
        assert wide.mean() > 0.5
        assert (g_hi[wide] >= 127).all()
    # canonical triangle vertices match the scene's index buffer
    v, idx, _, off = scene.arrays()
    ids = tris[:, 3].copy().view(np.uint32)
    order = np.argsort(ids)
    canon = tris[order].reshape(-1, 3, 4)[:, :, :3]
    inst = tris[order, 7].copy().view(np.uint32)
    prim = tris[order, 11].copy().view(np.uint32)
    vi = off[2 * inst + 1][:, None] + idx[(off[2 * inst] + 3 * prim)[:, None] + np.arange(3)]
    assert np.array_equal(canon, v["pos"][vi])


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH, N.XR_BUILDER_SAH_HOST])
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


def _leaf_sets(nodes: np.ndarray, tris: np.ndarray):
    """The tree's leaves as sorted tuples of canonical triangle ids, sorted."""
    _, _, links, counts = decode_nodes(nodes)
    ids = tris[:, 3].copy().view(np.uint32)
    out = []
    for i in range(len(nodes)):
        for k in range(int(counts[i])):
            if links[i, k] < 0:
                code = ~int(links[i, k])
                first, n = code >> 4, (code & 15) + 1
                out.append(tuple(sorted(ids[first:first + n].tolist())))
    return sorted(out)


@pytest.mark.parametrize("which", ["cornell", "terrain", "terrain_big"])
def test_gpu_sah_builds_the_host_sah_tree(which):
    """The GPU binned-SAH builder makes the host builder's decisions: the same
    leaves (as triangle sets), node count and depth, on scenes whose
    triangle centroids are distinct (the Cornell box, a 128^2 and a 512^2
    terrain with spheres: ranges above and below the GPU's 512-triangle
    one-wave threshold)."""
    scene = {"cornell": cornell_scene, "terrain": small_terrain,
             "terrain_big": lambda: N.Scene.terrain(3, 512, 128, 16)}[which]()
    say(f"[bvh] GPU vs host SAH on {which} ({scene.triangle_count} tris)")
    ctx = N.Context(scene, 8, 8, builder=N.XR_BUILDER_SAH)
    gpu = ctx.read_accel()
    g_info = ctx.accel_info()
    ctx.build_accel(N.XR_BUILDER_SAH_HOST)
    host = ctx.read_accel()
    h_info = ctx.accel_info()
    assert g_info.max_depth == h_info.max_depth and len(gpu[0]) == len(host[0])
    assert _leaf_sets(*gpu) == _leaf_sets(*host)
    say(f"[bvh] build ms GPU {g_info.build_ms:.1f} host {h_info.build_ms:.1f}")


def _degenerate_scene() -> N.Scene:
    """3,000 copies of one triangle (one centroid: the builders' median
    split), 1,000 zero-area triangles (points and segments), 1,000 triangles
    flat in y on one plane, 3,000 random small triangles in the unit box."""
    rng = np.random.default_rng(21)
    same = np.repeat(np.array([[[0.4, 0.4, 0.4], [0.6, 0.4, 0.4], [0.4, 0.6, 0.45]]]), 3000, axis=0)
    p = rng.uniform(0.0, 1.0, size=(1000, 1, 3))
    seg = p + rng.uniform(0.0, 0.01, size=(1000, 1, 3)) * np.array([[0.0], [1.0], [1.0]])[None, :, :]
    seg[:500] = p[:500]  # the first half are points
    flat = rng.uniform(0.0, 1.0, size=(1000, 3, 3))
    flat[:, :, 1] = 0.25
    small = rng.uniform(0.0, 1.0, size=(3000, 1, 3)) + rng.uniform(-0.02, 0.02, size=(3000, 3, 3))
    t = np.concatenate([same, seg, flat, small]).astype(np.float32)
    v = np.zeros(3 * len(t), N.VERTEX_DTYPE)
    v["pos"] = t.reshape(-1, 3)
    v["normal"] = [0, 1, 0]
    m = np.zeros(1, N.MATERIAL_DTYPE)
    m["base_color"] = [0.5, 0.5, 0.5, 1]
    m["base_color_texture_index"] = -1
    return N.Scene.from_arrays(v, np.arange(len(v), dtype=np.uint32), m, np.array([0, 0], np.uint32))


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH, N.XR_BUILDER_SAH_HOST])
def test_bvh_degenerate_triangles(builder):
    """Coincident centroids, zero-area and axis-flat triangles: a valid tree
    (every triangle in one leaf, every box holding what is below it) and
    closest / any hits bit for bit against brute force over all triangles."""
    from tests.helpers import oracle_scene, random_rays, bytes_equal_rows, describe_mismatch
    scene = _degenerate_scene()
    say(f"[bvh] degenerate scene ({scene.triangle_count} tris), builder {builder}")
    ctx = N.Context(scene, 8, 8, builder=builder)
    nodes, tris = ctx.read_accel()
    check_bvh(nodes, tris, scene.triangle_count)
    orc = oracle_scene(scene)
    rays = random_rays(8000, -0.2, 1.2, seed=23, tmin=0.0, tmax=np.inf)
    gpu = ctx.trace(rays)
    ref = orc.trace32(rays, brute=True)
    eq = bytes_equal_rows(gpu, ref)
    assert eq.all(), describe_mismatch("closest hit, degenerate scene", gpu, ref, eq)
    assert (gpu["t"] > 0).mean() > 0.1
    any_gpu = ctx.trace(rays, any_hit=True)
    assert np.array_equal(any_gpu["t"] >= 0, ref["t"] >= 0)
