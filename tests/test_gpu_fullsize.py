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


def test_fullsize_frames_bit_exact_on_pixel_subset(big):
    from oracle.binding import temporal as orc_temporal
    scene, cam, orc = big
    ctx = N.Context(scene, W, H)
    info = ctx.accel_info()
    say(f"[fullsize] BVH depth bound {info.max_depth}, {info.triangle_count} tris")
    params = N.default_params(visibility=N.XR_VIS_TRACE, temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0)
    rng = np.random.default_rng(11)
    px = np.sort(rng.choice(W * H, size=20000, replace=False)).astype(np.uint32)
    prev = cam.view(0)
    for k in range(3):
        frame = k + 1
        view = cam.view(k)
        ubo = N.make_ubo(view, prev, 45.0, W, H, bounces=4, frame=frame)
        prev = view
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
        orc_temporal(W, H, frame, 20, samples, motion, gpos, temporal, spatial, output, pixels=px)
        gt = ctx.read(N.XR_BUF_TEMPORAL)
        for slot in range(2):
            idx = px + slot * W * H
            eq = bytes_equal_rows(gt[idx], temporal[idx])
            assert eq.all(), describe_mismatch("temporal", gt[idx], temporal[idx], eq)
        p.passes = N.XR_PASS_SPATIAL
        gs_before = ctx.read(N.XR_BUF_SPATIAL).copy()
        ctx.render(ubo, N.XR_MODE_RESTIR, p)
        ctx.sync()
        orc.spatial(W, H, frame, 1, 5, 30.0, 300, 1, samples, gt, gs_before, output, pixels=px)
        gs = ctx.read(N.XR_BUF_SPATIAL)
        idx = px + (1 - frame % 2) * W * H
        eq = bytes_equal_rows(gs[idx], gs_before[idx])
        assert eq.all(), describe_mismatch("spatial", gs[idx], gs_before[idx], eq)


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
