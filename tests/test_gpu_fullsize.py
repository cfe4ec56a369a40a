"""Parity at the benchmark's full size (1920x1080, 2,748,546 triangles) on a
random pixel subset: the oracle restates the initial-sample, temporal and
spatial passes for those pixels from the GPU's own inputs and must agree bit
for bit.  Also a deep-BVH closest-hit batch against brute force (the
traversal stack spills past its LDS levels only on trees this deep)."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import bytes_equal_rows, describe_mismatch, oracle_scene, say
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


@pytest.fixture(scope="module")
def big():
    say("[fullsize] building terrain 1024^2 + 512 spheres + 64 lights")
    scene = N.Scene.terrain(1, 1024, 512, 64)
    cam = N.Camera.orbit(2, 8, 128)
    orc = oracle_scene(scene)
    return scene, cam, orc


PARAMS = dict(visibility=N.XR_VIS_TRACE, temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0)


def _check_frame_passes(ctx, orc, ubo, frame, px, params):
    """Renders one frame pass by pass; after each pass the oracle restates it
    for the pixels px from the GPU's own inputs and must agree bit for bit."""
    from oracle.binding import temporal as orc_temporal

    say(f"[fullsize] frame {frame}: initial pass")
    p = N.XrRestirParams.from_buffer_copy(params)
    p.passes = N.XR_PASS_INITIAL
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    samples = ctx.read(N.XR_BUF_SAMPLES).copy()
    valid = (samples["material_id"] != 0xFFFFFFFF).mean()
    say(f"[fullsize] valid primary hits {valid:.3f}")
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, pixels=px)
    eq = bytes_equal_rows(samples[px], ref["samples"][px])
    assert eq.all(), describe_mismatch("samples", samples[px], ref["samples"][px], eq)
    assert valid > 0.5
    motion = ctx.read(N.XR_BUF_MOTION).copy()
    gpos = np.ascontiguousarray(ctx.read(N.XR_BUF_GBUF_POS))
    assert np.array_equal(motion[px], ref["motion"][px])
    temporal = ctx.read(N.XR_BUF_TEMPORAL).copy()
    spatial = ctx.read(N.XR_BUF_SPATIAL).copy()
    output = ctx.read(N.XR_BUF_OUTPUT).copy()
    say(f"[fullsize] frame {frame}: temporal + spatial")
    p.passes = N.XR_PASS_TEMPORAL
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, temporal, spatial, output, pixels=px)
    gt = ctx.read(N.XR_BUF_TEMPORAL)
    for slot in range(2):
        idx = px + slot * W * H
        eq = bytes_equal_rows(gt[idx], temporal[idx])
        assert eq.all(), describe_mismatch("temporal", gt[idx], temporal[idx], eq)
    p.passes = N.XR_PASS_SPATIAL
    gs_before = ctx.read(N.XR_BUF_SPATIAL).copy()
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    orc.spatial(W, H, frame, 1, params.spatial_samples, params.spatial_radius, params.spatial_m_cap, 1, samples,
                gt, gs_before, output, pixels=px)
    gs = ctx.read(N.XR_BUF_SPATIAL)
    idx = px + (1 - frame % 2) * W * H
    eq = bytes_equal_rows(gs[idx], gs_before[idx])
    assert eq.all(), describe_mismatch("spatial", gs[idx], gs_before[idx], eq)


def test_fullsize_frames_bit_exact_on_pixel_subset(big):
    scene, cam, orc = big
    ctx = N.Context(scene, W, H)
    info = ctx.accel_info()
    say(f"[fullsize] BVH depth bound {info.max_depth}, {info.triangle_count} tris")
    params = N.default_params(**PARAMS)
    rng = np.random.default_rng(11)
    px = np.sort(rng.choice(W * H, size=20000, replace=False)).astype(np.uint32)
    prev = cam.view(0)
    for k in range(3):
        view = cam.view(k)
        _check_frame_passes(ctx, orc, N.make_ubo(view, prev, 45.0, W, H, bounces=4, frame=k + 1), k + 1, px, params)
        prev = view


def _check_reservoir_invariants(ctx, params, frame):
    """Properties every frame keeps, whatever its history: reservoir weights
    finite and non-negative, M within its pass's cap (Temporal.rgen:114-129,
    Spatial.rgen:149), every output pixel opaque."""
    for which, cap in ((N.XR_BUF_TEMPORAL, params.temporal_m_cap), (N.XR_BUF_SPATIAL, params.spatial_m_cap)):
        r = ctx.read(which)
        for f in ("w_sum", "w"):
            assert np.isfinite(r[f]).all() and (r[f] >= 0).all(), f"buffer {which} {f} at frame {frame}"
        assert (r["m"] <= cap).all(), f"buffer {which} M above {cap} at frame {frame}"
        say(f"[fullsize] frame {frame}: buffer {which} max M {int(r['m'].max())}, "
            f"W > 0 on {(r['w'] > 0).mean():.3f}")
    out = ctx.read(N.XR_BUF_OUTPUT)
    assert ((out >> 24) == 0xFF).all(), f"output alpha at frame {frame}"


def test_fullsize_long_history_on_the_camera_path(big):
    """The benchmark's 128-frame camera path at full size: after 63 and 127
    whole frames the reservoirs keep their invariants with M saturated at the
    caps, and frame 128 still equals the oracle pass by pass on a pixel
    subset (long histories: capped M, rescaled w_sum, reprojection across
    the path's motion)."""
    scene, cam, orc = big
    ctx = N.Context(scene, W, H, builder=N.XR_BUILDER_SAH)
    params = N.default_params(**PARAMS)
    frames = 128
    views = [cam.view(k) for k in range(frames)]
    for k in range(frames - 1):
        ubo = N.make_ubo(views[k], views[k - 1] if k else views[k], 45.0, W, H, bounces=4, frame=k + 1)
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        if k + 1 in (63, frames - 1):
            ctx.sync()
            _check_reservoir_invariants(ctx, params, k + 1)
    t = ctx.read(N.XR_BUF_TEMPORAL)
    saturated = (t["m"] == params.temporal_m_cap).mean()
    say(f"[fullsize] temporal M at its cap on {saturated:.3f} of the reservoirs")
    assert saturated > 0.3, "the history reached the temporal M cap"
    rng = np.random.default_rng(12)
    px = np.sort(rng.choice(W * H, size=20000, replace=False)).astype(np.uint32)
    ubo = N.make_ubo(views[-1], views[-2], 45.0, W, H, bounces=4, frame=frames)
    _check_frame_passes(ctx, orc, ubo, frames, px, params)
    _check_reservoir_invariants(ctx, params, frames)


def test_deep_bvh_trace_matches_oracle(big):
    scene, cam, orc = big
    ctx = N.Context(scene, 8, 8)
    rng = np.random.default_rng(5)
    n = 100000
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = np.stack([rng.uniform(0, 1, n), rng.uniform(-0.05, 0.6, n), rng.uniform(0, 1, n)], 1)
    d = rng.normal(size=(n, 3))
    rays["direction"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, 1e4
    gpu = ctx.trace(rays)
    ref = orc.trace32(rays)
    eq = bytes_equal_rows(gpu, ref)
    assert eq.all(), describe_mismatch("deep closest hit", gpu, ref, eq)


@pytest.mark.parametrize("fp8", [False, True])
def test_fullsize_cache_long_run_stays_finite(big, fp8):
    """128 cache frames along the camera path at full size (training on 65,536
    samples per frame, no spatial neighbours, the equal-error setting of the
    benchmark's study; FP8 weights in the second case): the online-trained
    network's parameters, the radiance behind the image and the reservoirs
    stay finite, and the training step keeps choosing its full batch."""
    scene, cam, orc = big
    ctx = N.Context(scene, W, H, builder=N.XR_BUILDER_SAH)
    ctx.mlp_reset(1)
    flags = N.XR_FLAG_MLP_CACHE | N.XR_FLAG_RADIANCE | (N.XR_FLAG_MLP_FP8 if fp8 else 0)
    params = N.default_params(**dict(PARAMS, spatial_samples=0), flags=flags, mlp_spread=0.01)
    frames = 128
    views = [cam.view(k) for k in range(frames)]
    w0 = ctx.mlp_params().copy()
    ctx.reset_counters()
    for k in range(frames):
        ubo = N.make_ubo(views[k], views[k - 1] if k else views[k], 45.0, W, H, bounces=4, frame=k + 1)
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        if k + 1 in (32, frames):
            ctx.sync()
            w = ctx.mlp_params()
            assert np.isfinite(w).all(), f"network parameters at frame {k + 1}"
            rad = ctx.read(N.XR_BUF_RADIANCE)
            assert np.isfinite(rad).all(), f"radiance at frame {k + 1}"
            _check_reservoir_invariants(ctx, params, k + 1)
            say(f"[fullsize] cache frame {k + 1} ({'FP8' if fp8 else 'BF16'}): max |param| "
                f"{np.abs(w).max():.3g}, mean radiance {rad[:, :3].mean():.4f}")
    assert not np.array_equal(ctx.mlp_params(), w0), "the network trained"
    queries, trained = ctx.counters_all()[3:]
    say(f"[fullsize] {queries / frames:.0f} cache queries and {trained / frames:.0f} training samples per frame")
    assert trained == frames * params.mlp_train_pixels
    assert queries > 0
