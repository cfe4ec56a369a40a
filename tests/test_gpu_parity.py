"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

The kernels and the oracle evaluate the same FP32 expressions in the same
order (no contraction, correctly rounded division and sqrt, the same cos/sin
polynomial), so traversal results, samples, G-buffers and reservoirs are
compared bit for bit.  The tone-mapped BGRA8 image goes through powf, whose
last bit may differ between the device library and glibc, so it is compared
to 1 LSB per channel.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests.helpers import (CORNELL_EYE, CORNELL_TARGET, bytes_equal_rows, cornell_scene, describe_mismatch,
                           look_view, oracle_scene, random_rays, rays_f64, say, small_terrain)
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _max_lsb_diff(a: np.ndarray, b: np.ndarray) -> int:
    a8 = a.view(np.uint8).astype(np.int32)
    b8 = b.view(np.uint8).astype(np.int32)
    return int(np.abs(a8 - b8).max()) if len(a8) else 0


# ---------------------------------------------------------------------------
# TraceRay
# ---------------------------------------------------------------------------
BUILDERS = [N.XR_BUILDER_LBVH, N.XR_BUILDER_PLOC, N.XR_BUILDER_SAH, N.XR_BUILDER_SAH_HOST]


@pytest.mark.parametrize("builder", BUILDERS)
@pytest.mark.parametrize("which", ["cornell", "terrain"])
def test_trace_closest_bit_exact(which, builder):
    scene = cornell_scene() if which == "cornell" else small_terrain()
    say(f"[trace] {which}: {scene.triangle_count} tris, builder {builder}")
    ctx = N.Context(scene, 8, 8, builder=builder)
    orc = oracle_scene(scene)
    lo, hi = ((0.02, 0.98) if which == "cornell" else (0.0, 0.5))
    rays = random_rays(20000, lo, hi, seed=7)
    gpu = ctx.trace(rays)
    ref = orc.trace32(rays)
    eq = bytes_equal_rows(gpu, ref)
    assert eq.all(), describe_mismatch("closest hit", gpu, ref, eq)
    assert (gpu["t"] > 0).mean() > 0.3  # the batch actually hits geometry


@pytest.mark.parametrize("builder", BUILDERS)
def test_trace_matches_fp64_reference(builder):
    """Criterion: >= 99.99% identical closest-hit ids vs an FP64
    search; every mismatch is a distance tie within 1e-6 * t."""
    scene = small_terrain(grid=256, spheres=64, lights=16)
    ctx = N.Context(scene, 8, 8, builder=builder)
    orc = oracle_scene(scene)
    rng = np.random.default_rng(3)
    n = 50000
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = np.stack([rng.uniform(0, 1, n), rng.uniform(0.3, 0.6, n), rng.uniform(0, 1, n)], 1)
    d = np.stack([rng.normal(size=n), -np.abs(rng.normal(size=n)) - 0.2, rng.normal(size=n)], 1)
    rays["direction"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = 1e4
    say(f"[trace-fp64] {scene.triangle_count} tris, {n} rays")
    gpu = ctx.trace(rays)
    ids, ts = orc.trace64(rays_f64(rays))
    same = gpu["triangle"] == ids
    assert same.mean() >= 0.9999, f"id match {same.mean():.6f}"
    bad = ~same
    if bad.any():
        t_gpu = gpu["t"][bad].astype(np.float64)
        t_ref = ts[bad]
        assert np.all(np.abs(t_gpu - t_ref) <= 1e-6 * np.maximum(np.abs(t_ref), 1e-3) + 1e-7), \
            list(zip(t_gpu[:5], t_ref[:5]))


@pytest.mark.parametrize("builder", BUILDERS)
def test_trace_any_hit_matches(builder):
    scene = small_terrain()
    ctx = N.Context(scene, 8, 8, builder=builder)
    orc = oracle_scene(scene)
    rays = random_rays(20000, 0.0, 0.6, seed=11, tmin=0.0, tmax=0.2)
    gpu = ctx.trace(rays, any_hit=True)
    ref = orc.trace32(rays, any_hit=True)
    assert np.array_equal(gpu["t"] >= 0, ref["t"] >= 0)


@pytest.mark.parametrize("builder", [N.XR_BUILDER_LBVH, N.XR_BUILDER_SAH])
@pytest.mark.parametrize("dist", [3.0, 100.0, 1000.0])
def test_trace_far_origins_match_brute_force(dist, builder):
    """Rays from far outside the scene (dist scene extents from its centre)
    aimed at random points of its box, tmax +inf: closest hits bit for bit
    against brute force over every triangle, any-hit flags equal.  Far
    origins stress the slab test's relative tolerance and the empty boxes of
    unused child slots (entered only by rays from beyond ~64 extents, where
    they lead to a real triangle's test)."""
    scene = cornell_scene()
    ctx = N.Context(scene, 8, 8, builder=builder)
    orc = oracle_scene(scene)
    rng = np.random.default_rng(int(dist) + 17)
    n = 6000
    centre = np.array([0.5, 0.5, 0.5])
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    org = (centre + dist * u).astype(np.float32)
    aim = rng.uniform(0.0, 1.0, size=(n, 3))
    d = aim - org.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = org
    rays["direction"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    say(f"[trace-far] origins {dist} extents away, builder {builder}")
    gpu = ctx.trace(rays)
    ref = orc.trace32(rays, brute=True)
    eq = bytes_equal_rows(gpu, ref)
    assert eq.all(), describe_mismatch("closest hit, far origins", gpu, ref, eq)
    assert (gpu["t"] > 0).mean() > 0.5
    any_gpu = ctx.trace(rays, any_hit=True)
    assert np.array_equal(any_gpu["t"] >= 0, ref["t"] >= 0)


def _mixed_scale_scene() -> N.Scene:
    """A 2000-unit ground quad, 2000 millimetre triangles in a unit box far from
    the origin, 200 triangles flat in x (zero extent on one axis) in the same
    box and one large far triangle: node grids of very different steps."""
    rng = np.random.default_rng(5)
    c = np.array([500.0, 0.5, 500.0])
    tris = [np.array([[-1000, 0, -1000], [1000, 0, -1000], [1000, 0, 1000]], np.float64),
            np.array([[-1000, 0, -1000], [1000, 0, 1000], [-1000, 0, 1000]], np.float64),
            np.array([[-900, 40, -900], [-800, 60, -900], [-900, 50, -800]], np.float64)]
    base = c + rng.uniform(-0.5, 0.5, size=(2000, 3))
    tris += list(base[:, None, :] + rng.uniform(-1e-3, 1e-3, size=(2000, 3, 3)))
    flat = c + rng.uniform(-0.5, 0.5, size=(200, 3))
    f = flat[:, None, :] + rng.uniform(-0.05, 0.05, size=(200, 3, 3))
    f[:, :, 0] = 500.2
    tris += list(f)
    t = np.array(tris, np.float32)
    v = np.zeros(3 * len(t), N.VERTEX_DTYPE)
    v["pos"] = t.reshape(-1, 3)
    v["normal"] = [0, 1, 0]
    m = np.zeros(1, N.MATERIAL_DTYPE)
    m["base_color"] = [0.5, 0.5, 0.5, 1]
    m["base_color_texture_index"] = -1
    return N.Scene.from_arrays(v, np.arange(len(v), dtype=np.uint32), m, np.array([0, 0], np.uint32))


@pytest.mark.parametrize("builder", BUILDERS)
def test_trace_mixed_scales_axis_rays_infinite_tmax(builder):
    """Closest hits against brute force over every triangle, and any-hit
    flags, for rays from inside the millimetre cluster: random, axis-parallel
    (two zero direction components), one zero component; tmin 0 or 1e-4,
    tmax 1e4 or +inf."""
    scene = _mixed_scale_scene()
    ctx = N.Context(scene, 8, 8, builder=builder)
    orc = oracle_scene(scene)
    rng = np.random.default_rng(9)
    n = 4000
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = (np.array([500.0, 0.5, 500.0]) + rng.uniform(-0.6, 0.6, size=(n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    q = n // 4
    axis = rng.integers(0, 3, size=q)
    d[q:2 * q] = 0.0
    d[np.arange(q, 2 * q), axis] = rng.choice([-1.0, 1.0], size=q)
    d[np.arange(2 * q, 3 * q), rng.integers(0, 3, size=q)] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["direction"] = d.astype(np.float32)
    rays["tmin"] = np.where(rng.random(n) < 0.5, 0.0, 1e-4)
    rays["tmax"] = np.where(rng.random(n) < 0.5, np.inf, 1e4)
    say(f"[trace-mixed] {scene.triangle_count} tris, builder {builder}")
    gpu = ctx.trace(rays)
    ref = orc.trace32(rays, brute=True)
    eq = bytes_equal_rows(gpu, ref)
    assert eq.all(), describe_mismatch("closest hit", gpu, ref, eq)
    assert 0.2 < (gpu["t"] > 0).mean() < 0.95
    seg = rays.copy()
    seg["tmax"] = np.where(np.isinf(seg["tmax"]), np.inf, 0.05)
    gany = ctx.trace(seg, any_hit=True)
    rany = orc.trace32(seg, any_hit=True)
    assert np.array_equal(gany["t"] >= 0, rany["t"] >= 0)


# ---------------------------------------------------------------------------
# ReSTIR GI frame, pass by pass
# ---------------------------------------------------------------------------
def _frames(scene, W, H, fov, eyes, target, bounces, params, check, sky=0):
    """Render frames pass by pass on the GPU; after each pass compare the
    pass's outputs with the oracle run on the GPU's own inputs."""
    from oracle.binding import temporal as orc_temporal
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    views = [look_view(e, target) for e in eyes]
    for k, view in enumerate(views):
        frame = k + 1
        ubo = N.make_ubo(view, views[k - 1] if k else view, fov, W, H, spp=1, total=1, bounces=bounces, sky=sky,
                         reuse=1, frame=frame)
        ub = N.ubo_to_bytes(ubo)
        say(f"[restir] frame {frame} {W}x{H} bounces={bounces} vis={params.visibility}")
        # initial sample pass
        p = N.XrRestirParams.from_buffer_copy(params)
        p.passes = N.XR_PASS_INITIAL
        ctx.render(ubo, N.XR_MODE_RESTIR, p)
        ctx.sync()
        ref = orc.initial_sample(ub, W, H)
        for name, which, key in [("samples", N.XR_BUF_SAMPLES, "samples"), ("gpos", N.XR_BUF_GBUF_POS, "gpos"),
                                 ("gnrm", N.XR_BUF_GBUF_NRM, "gnrm"), ("gmatuv", N.XR_BUF_GBUF_MATUV, "gmatuv"),
                                 ("motion", N.XR_BUF_MOTION, "motion")]:
            g = ctx.read(which)
            eq = bytes_equal_rows(g, ref[key])
            check(eq, name, g, ref[key])
        # temporal pass on the GPU's state
        samples = ctx.read(N.XR_BUF_SAMPLES)
        motion = ctx.read(N.XR_BUF_MOTION)
        gpos = np.ascontiguousarray(ctx.read(N.XR_BUF_GBUF_POS))
        temporal = ctx.read(N.XR_BUF_TEMPORAL).copy()
        spatial = ctx.read(N.XR_BUF_SPATIAL).copy()
        output = ctx.read(N.XR_BUF_OUTPUT).copy()
        p.passes = N.XR_PASS_TEMPORAL
        ctx.render(ubo, N.XR_MODE_RESTIR, p)
        ctx.sync()
        orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, temporal, spatial, output)
        gt = ctx.read(N.XR_BUF_TEMPORAL)
        check(bytes_equal_rows(gt, temporal), "temporal reservoirs", gt, temporal)
        gs = ctx.read(N.XR_BUF_SPATIAL)
        check(bytes_equal_rows(gs, spatial), "spatial reservoirs (temporal pass)", gs, spatial)
        go = ctx.read(N.XR_BUF_OUTPUT)
        assert _max_lsb_diff(go, output) <= 1, "temporal output"
        # spatial pass
        output = go.copy()
        p.passes = N.XR_PASS_SPATIAL
        ctx.render(ubo, N.XR_MODE_RESTIR, p)
        ctx.sync()
        orc.spatial(W, H, frame, 1, params.spatial_samples, params.spatial_radius, params.spatial_m_cap,
                    params.visibility, samples, gt, spatial, output)
        gs = ctx.read(N.XR_BUF_SPATIAL)
        check(bytes_equal_rows(gs, spatial), "spatial reservoirs", gs, spatial)
        go = ctx.read(N.XR_BUF_OUTPUT)
        assert _max_lsb_diff(go, output) <= 1, "spatial output"
    return ctx


def _strict(eq, name, g, r):
    assert eq.all(), describe_mismatch(name, g, r, eq)


@pytest.mark.parametrize("bounces", [4, 10])
def test_restir_cornell_reference_mode(bounces):
    scene = cornell_scene()
    eyes = [CORNELL_EYE, (0.52, 0.5, -1.2), (0.54, 0.51, -1.19)]
    params = N.default_params()
    _frames(scene, 96, 64, 60.0, eyes, CORNELL_TARGET, bounces, params, _strict)


def test_restir_cornell_traced_visibility_config_constants():
    scene = cornell_scene()
    eyes = [CORNELL_EYE, (0.51, 0.5, -1.2)]
    params = N.default_params(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)
    _frames(scene, 80, 60, 60.0, eyes, CORNELL_TARGET, 4, params, _strict)


@pytest.mark.parametrize("reuse", [1, 0])
def test_restir_without_spatial_neighbours(reuse):
    """spatial_samples = 0: one kernel runs the temporal pass and the spatial
    pass (which then reads nothing but the pixel's own temporal reservoir).
    Whole frames must equal the oracle's temporal pass followed by its
    spatial pass with no neighbours (Spatial.rgen:64-183 with
    NUM_SPATIAL_SAMPLES = 0); a spatial pass with neighbours run alone
    afterwards must rebuild the neighbour records the fused kernel skips."""
    from oracle.binding import temporal as orc_temporal
    scene = cornell_scene()
    W, H = 80, 60
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    eyes = [CORNELL_EYE, (0.51, 0.5, -1.2), (0.52, 0.5, -1.2)]
    views = [look_view(e, CORNELL_TARGET) for e in eyes]
    params = N.default_params(temporal_m_cap=20, spatial_samples=0, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)
    for k, view in enumerate(views):
        frame = k + 1
        ubo = N.make_ubo(view, views[k - 1] if k else view, 60.0, W, H, spp=1, total=1, bounces=4, reuse=reuse,
                         frame=frame)
        say(f"[restir] frame {frame} without spatial neighbours, reuse {reuse}")
        temporal = ctx.read(N.XR_BUF_TEMPORAL).copy()
        spatial = ctx.read(N.XR_BUF_SPATIAL).copy()
        output = ctx.read(N.XR_BUF_OUTPUT).copy()
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        ctx.sync()
        samples = ctx.read(N.XR_BUF_SAMPLES).copy()
        ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H)["samples"]
        _strict(bytes_equal_rows(samples, ref), "samples", samples, ref)
        motion = ctx.read(N.XR_BUF_MOTION).copy()
        gpos = np.ascontiguousarray(ctx.read(N.XR_BUF_GBUF_POS))
        orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, temporal, spatial, output)
        orc.spatial(W, H, frame, reuse, 0, params.spatial_radius, params.spatial_m_cap, params.visibility, samples,
                    temporal, spatial, output)
        gt = ctx.read(N.XR_BUF_TEMPORAL)
        _strict(bytes_equal_rows(gt, temporal), "temporal reservoirs", gt, temporal)
        gs = ctx.read(N.XR_BUF_SPATIAL)
        _strict(bytes_equal_rows(gs, spatial), "spatial reservoirs", gs, spatial)
        go = ctx.read(N.XR_BUF_OUTPUT)
        assert _max_lsb_diff(go, output) <= 1, "output"
        assert (go != 0).any()
    # the last frame's spatial pass again, with neighbours, on its own
    p = N.XrRestirParams.from_buffer_copy(params)
    p.spatial_samples = 3
    p.passes = N.XR_PASS_SPATIAL
    gt = ctx.read(N.XR_BUF_TEMPORAL).copy()
    spatial = ctx.read(N.XR_BUF_SPATIAL).copy()
    output = ctx.read(N.XR_BUF_OUTPUT).copy()
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    orc.spatial(W, H, len(views), reuse, 3, p.spatial_radius, p.spatial_m_cap, p.visibility, samples, gt, spatial,
                output)
    gs = ctx.read(N.XR_BUF_SPATIAL)
    _strict(bytes_equal_rows(gs, spatial), "spatial reservoirs with neighbours", gs, spatial)
    assert _max_lsb_diff(ctx.read(N.XR_BUF_OUTPUT), output) <= 1
    ctx.close()


def test_restir_terrain_sky():
    scene = small_terrain()
    eyes = [(0.2, 0.4, 0.2), (0.21, 0.4, 0.2)]
    params = N.default_params(visibility=N.XR_VIS_TRACE)
    _frames(scene, 64, 48, 45.0, eyes, (0.5, 0.0, 0.5), 4, params, _strict, sky=1)


# The reference's own cb2 scene (asset/model/cb2, CC-BY-4.0, see
# tests/golden/gltf/cb2/license.txt) through the glTF loader, at the
# reference's default view (Scene.cpp:29), fov and bounce count (main.cpp:45-47).
CB2 = Path(__file__).parent / "golden" / "gltf" / "cb2" / "scene.gltf"
CB2_EYE, CB2_TARGET = (0.0, 1.0, 3.0), (0.0, 1.0, 0.0)


@pytest.mark.parametrize("vis", [N.XR_VIS_REFERENCE, N.XR_VIS_TRACE])
def test_restir_cb2_gltf(vis):
    scene = N.Scene.gltf(CB2)
    eyes = [CB2_EYE, (0.02, 1.0, 3.0), (0.03, 1.01, 2.99)]
    params = N.default_params(visibility=vis)
    _frames(scene, 96, 64, 45.0, eyes, CB2_TARGET, 10, params, _strict)


SB2 = Path(__file__).parent / "golden" / "gltf" / "sb2" / "scene.gltf"


def test_restir_sb2_gltf_textured():
    """The reference's sb2 scene (75,277 triangles, PNG textures) at its default view."""
    scene = N.Scene.gltf(SB2)
    assert scene.textures()
    eyes = [CB2_EYE, (0.05, 1.0, 3.0)]
    _frames(scene, 64, 48, 45.0, eyes, CB2_TARGET, 10, N.default_params(visibility=N.XR_VIS_TRACE), _strict)


def test_plain_cb2_gltf(tmp_path):
    scene = N.Scene.gltf(CB2)
    W, H = 80, 48
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    view = look_view(CB2_EYE, CB2_TARGET)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    for frame in range(1, 3):
        ubo = N.make_ubo(view, view, 45.0, W, H, spp=1, total=frame, bounces=10, sky=0, reuse=0, frame=frame)
        say(f"[plain cb2] frame {frame}")
        ctx.render(ubo, N.XR_MODE_PLAIN, N.default_params())
        ctx.sync()
        orc.plain(N.ubo_to_bytes(ubo), W, H, 0, accum, out)
        ga = ctx.read(N.XR_BUF_ACCUMULATION)
        eq = bytes_equal_rows(ga, accum)
        assert eq.all(), describe_mismatch("accumulation", ga, accum, eq)
        assert _max_lsb_diff(ctx.read(N.XR_BUF_OUTPUT), out) <= 1
    assert accum[:, :3].max() > 0
    # the image dump holds the output image's R, G, B bytes
    import zlib
    ctx.save_png(tmp_path / "cb2.png")
    png = (tmp_path / "cb2.png").read_bytes()
    idat = png[8 + 25 + 8:]  # IHDR chunk is 25 bytes; IDAT data follows its 8-byte header
    raw = np.frombuffer(zlib.decompress(idat[:int.from_bytes(png[33:37], "big")]), np.uint8)
    rgb = raw.reshape(H, 1 + 3 * W)[:, 1:].reshape(H * W, 3)
    img = ctx.read_image()
    assert np.array_equal(rgb[:, 0], (img >> 16) & 0xFF) and np.array_equal(rgb[:, 2], img & 0xFF)


# Textured materials: the Cornell scene with planar texture coordinates that
# run past [0, 1] (clamp to edge), alternate diffuse materials on a 37x23 or
# a 5x9 noise texture and the emitter naming a 1x1 texture, which the bounce
# loops never sample: emission stays baseColor * 4 (InitialSample.rgen:117-120,
# rgen.hlsl:96-99).
def textured_cornell(emitter_texture: bool = True) -> N.Scene:
    v, i, m, o = cornell_scene().arrays()
    v, m = v.copy(), m.copy()
    v["tex_coord"][:, 0] = v["pos"][:, 0] * 1.7 - 0.2
    v["tex_coord"][:, 1] = v["pos"][:, 1] * 1.3 + v["pos"][:, 2] * 0.5 - 0.1
    rng = np.random.default_rng(5)
    tex = [(37, 23, rng.integers(0, 256, 37 * 23 * 4, dtype=np.uint8)),
           (5, 9, rng.integers(0, 256, 5 * 9 * 4, dtype=np.uint8)),
           (1, 1, np.array([200, 100, 50, 255], np.uint8))]
    lam = np.flatnonzero(m["material_model"] == N.MATERIAL_LAMBERTIAN)
    m["base_color_texture_index"][lam[::2]] = 0
    m["base_color_texture_index"][lam[1::2]] = 1
    m["base_color_texture_index"][m["material_model"] == N.MATERIAL_DIFFUSE_LIGHT] = 2 if emitter_texture else -1
    return N.Scene.from_arrays(v, i, m, o, tex)


@pytest.mark.parametrize("vis", [N.XR_VIS_REFERENCE, N.XR_VIS_TRACE])
def test_restir_textured(vis):
    eyes = [CORNELL_EYE, (0.52, 0.5, -1.2)]
    params = N.default_params(visibility=vis, spatial_samples=5, spatial_radius=30.0)
    _frames(textured_cornell(), 72, 56, 60.0, eyes, CORNELL_TARGET, 4, params, _strict)


def test_plain_textured():
    scene = textured_cornell()
    W, H = 64, 48
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    for frame in range(1, 3):
        ubo = N.make_ubo(view, view, 60.0, W, H, spp=1, total=frame, bounces=10, sky=0, reuse=0, frame=frame)
        say(f"[plain textured] frame {frame}")
        ctx.render(ubo, N.XR_MODE_PLAIN, N.default_params(plain_max_depth=0))
        ctx.sync()
        orc.plain(N.ubo_to_bytes(ubo), W, H, 0, accum, out)
        ga = ctx.read(N.XR_BUF_ACCUMULATION)
        eq = bytes_equal_rows(ga, accum)
        assert eq.all(), describe_mismatch("accumulation", ga, accum, eq)
    assert accum[:, :3].max() > 0
    # the emitter's texture is never sampled: the same frames without it are identical
    ctx2 = N.Context(textured_cornell(emitter_texture=False), W, H)
    for frame in range(1, 3):
        ubo = N.make_ubo(view, view, 60.0, W, H, spp=1, total=frame, bounces=10, sky=0, reuse=0, frame=frame)
        ctx2.render(ubo, N.XR_MODE_PLAIN, N.default_params(plain_max_depth=0))
        ctx2.sync()
    assert ctx2.read(N.XR_BUF_ACCUMULATION).tobytes() == ctx.read(N.XR_BUF_ACCUMULATION).tobytes()


def test_missing_texture_refuses_the_frame():
    v, i, m, o = cornell_scene().arrays()
    m = m.copy()
    m["base_color_texture_index"][0] = 1
    scene = N.Scene.from_arrays(v, i, m, o, [(1, 1, np.full(4, 255, np.uint8))])  # only texture 0
    ctx = N.Context(scene, 16, 16)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    ubo = N.make_ubo(view, view, 60.0, 16, 16)
    with pytest.raises(N.XrError) as e:
        ctx.render(ubo, N.XR_MODE_RESTIR, N.default_params())
    assert e.value.code == N.XR_ERR_STATE
    ctx.set_textures([(1, 1, np.full(4, 255, np.uint8)), (2, 1, np.full(8, 9, np.uint8))])
    ctx.render(ubo, N.XR_MODE_RESTIR, N.default_params())
    ctx.sync()


# ---------------------------------------------------------------------------
# Plain path tracer
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("max_depth", [3, 5, 0])
@pytest.mark.parametrize("spp", [1, 2])
def test_plain_accumulation(max_depth, spp):
    """Plain path tracing over three accumulated frames: one sample per pixel
    at unbounded depth runs as a head kernel plus one launch per bounce (a
    closed box: most paths survive, so the launches hand paths on); two
    samples, or depth 3, in one kernel."""
    scene = cornell_scene()
    W, H = 80, 48
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    params = N.default_params(plain_max_depth=max_depth)
    for frame in range(1, 4):
        ubo = N.make_ubo(view, view, 60.0, W, H, spp=spp, total=frame * spp, bounces=10, sky=0, reuse=0, frame=frame)
        say(f"[plain] frame {frame} depth={max_depth} spp={spp}")
        ctx.render(ubo, N.XR_MODE_PLAIN, params)
        ctx.sync()
        orc.plain(N.ubo_to_bytes(ubo), W, H, max_depth, accum, out)
        ga = ctx.read(N.XR_BUF_ACCUMULATION)
        eq = bytes_equal_rows(ga, accum)
        assert eq.all(), describe_mismatch("accumulation", ga, accum, eq)
        go = ctx.read(N.XR_BUF_OUTPUT)
        assert _max_lsb_diff(go, out) <= 1
    assert accum[:, :3].max() > 0  # the light is visible / reached


def test_rays_counted():
    scene = cornell_scene()
    W, H = 64, 64
    ctx = N.Context(scene, W, H)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    ctx.reset_counters()
    ctx.render(ubo, N.XR_MODE_RESTIR, N.default_params(visibility=N.XR_VIS_TRACE))
    ctx.sync()
    ext, shadow = ctx.counters()
    assert W * H <= ext <= 4 * W * H
    assert 0 < shadow <= 6 * W * H
    t = ctx.pass_times()
    assert len(t) == 4 and all(x >= 0 for x in t)  # initial, temporal, spatial, cache
    assert t[3] < 0.05  # no cache work without the cache flag (two adjacent events)


# ---------------------------------------------------------------------------
# Image sizes that fill no tile: single pixels, single rows and columns, sizes
# that are no multiple of the 32 x 8 tiles
# ---------------------------------------------------------------------------
ODD_SIZES = [(1, 1), (1, 9), (9, 1), (33, 9), (37, 11)]


@pytest.mark.parametrize("size", ODD_SIZES)
def test_restir_odd_sizes(size):
    """ReSTIR GI pass by pass at the configuration's constants (traced
    visibility, 5 neighbours within 30 px, which reach past every edge of
    these images) against the oracle, two frames."""
    W, H = size
    scene = cornell_scene()
    eyes = [CORNELL_EYE, (0.51, 0.5, -1.2)]
    params = N.default_params(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)
    _frames(scene, W, H, 60.0, eyes, CORNELL_TARGET, 4, params, _strict)


@pytest.mark.parametrize("size", ODD_SIZES)
def test_restir_whole_frames_odd_sizes(size):
    """Whole frames in one render call (the benchmark's form) equal the pass-by-
    pass frames: same samples, reservoirs and output bytes after three frames."""
    W, H = size
    scene = cornell_scene()
    eyes = [CORNELL_EYE, (0.51, 0.5, -1.2), (0.52, 0.5, -1.2)]
    views = [look_view(e, CORNELL_TARGET) for e in eyes]
    params = N.default_params(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)
    whole = N.Context(scene, W, H)
    split = N.Context(scene, W, H)
    for k, view in enumerate(views):
        ubo = N.make_ubo(view, views[k - 1] if k else view, 60.0, W, H, bounces=4, reuse=1, frame=k + 1)
        whole.render(ubo, N.XR_MODE_RESTIR, params)
        for passes in (N.XR_PASS_INITIAL, N.XR_PASS_TEMPORAL, N.XR_PASS_SPATIAL):
            p = N.XrRestirParams.from_buffer_copy(params)
            p.passes = passes
            split.render(ubo, N.XR_MODE_RESTIR, p)
    whole.sync()
    split.sync()
    for which in (N.XR_BUF_SAMPLES, N.XR_BUF_TEMPORAL, N.XR_BUF_SPATIAL, N.XR_BUF_OUTPUT):
        a, b = whole.read(which), split.read(which)
        assert a.tobytes() == b.tobytes(), f"buffer {which} differs at {W}x{H}"


@pytest.mark.parametrize("size", [(1, 1), (9, 1), (33, 9)])
@pytest.mark.parametrize("max_depth", [3, 0])
def test_plain_odd_sizes(size, max_depth):
    """Plain accumulation over two frames, one kernel (depth 3) or the head plus
    per-bounce launches (unbounded depth), against the oracle."""
    W, H = size
    scene = cornell_scene()
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    params = N.default_params(plain_max_depth=max_depth)
    for frame in (1, 2):
        ubo = N.make_ubo(view, view, 60.0, W, H, spp=1, total=frame, bounces=10, sky=0, reuse=0, frame=frame)
        ctx.render(ubo, N.XR_MODE_PLAIN, params)
        ctx.sync()
        orc.plain(N.ubo_to_bytes(ubo), W, H, max_depth, accum, out)
        ga = ctx.read(N.XR_BUF_ACCUMULATION)
        eq = bytes_equal_rows(ga, accum)
        assert eq.all(), describe_mismatch(f"accumulation {W}x{H}", ga, accum, eq)
        assert _max_lsb_diff(ctx.read(N.XR_BUF_OUTPUT), out) <= 1
