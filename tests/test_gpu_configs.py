"""The larger configurations against the oracle at their own workloads, on
pixel subsets (needs an MI355X).

  * 12 batched views of uneven size (4x 256x256, 4x 960x540, 2x 1920x1080,
    2x 1280x720) of the 2.75 M-triangle terrain, each on its own seeded
    camera path, one launch per pass for all views, replayed as a HIP graph
    with the autotuner's kernel configuration.

After every frame each view's buffers are read back and the oracle restates,
for a random subset of that view's pixels, the initial sample (from the
UBO), the temporal pass (from the frame's samples, motion, G-buffer and the
previous frame's reservoirs, which the frame leaves in place) and the spatial
pass (from the frame's samples and temporal reservoirs); the results must
agree bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import bytes_equal_rows, describe_mismatch, oracle_scene, say
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu

C3_SIZES = [(256, 256)] * 4 + [(960, 540)] * 4 + [(1920, 1080)] * 2 + [(1280, 720)] * 2
PARAMS = dict(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, spatial_m_cap=300,
              visibility=N.XR_VIS_TRACE)


def _eq(name, got, ref):
    eq = bytes_equal_rows(got, ref)
    assert eq.all(), describe_mismatch(name, got, ref, eq)


def check_view_frame(orc, read, W, H, ubo, frame, px, params):
    """The three passes of one frame of one view on the pixels px: `read(which)`
    returns the view's buffer after the frame."""
    from oracle.binding import temporal as orc_temporal
    n = W * H
    samples = read(N.XR_BUF_SAMPLES).copy()
    motion = read(N.XR_BUF_MOTION).copy()
    gpos = np.ascontiguousarray(read(N.XR_BUF_GBUF_POS))
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, pixels=px)
    _eq("samples", samples[px], ref["samples"][px])
    assert np.array_equal(motion[px], ref["motion"][px]) and np.array_equal(gpos[px], ref["gpos"][px])
    t_gpu = read(N.XR_BUF_TEMPORAL).copy()
    s_gpu = read(N.XR_BUF_SPATIAL).copy()
    cur, wr = frame % 2, 1 - frame % 2
    # temporal: previous slot as the GPU left it; current slot restated
    t_ref = t_gpu.copy()
    s_ref = s_gpu.copy()
    orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, t_ref, s_ref, np.zeros(n, np.uint32),
                 pixels=px)
    _eq("temporal reservoirs", t_gpu[px + cur * n], t_ref[px + cur * n])
    _eq("spatial slot written by the temporal pass", s_gpu[px + cur * n], s_ref[px + cur * n])
    # spatial: from the frame's samples and temporal reservoirs
    s_ref = s_gpu.copy()
    orc.spatial(W, H, frame, 1, params.spatial_samples, params.spatial_radius, params.spatial_m_cap,
                params.visibility, samples, t_gpu, s_ref, np.zeros(n, np.uint32), pixels=px)
    _eq("spatial reservoirs", s_gpu[px + wr * n], s_ref[px + wr * n])
    return float((samples["material_id"] != 0xFFFFFFFF).mean())


@pytest.fixture(scope="module")
def terrain():
    say("[configs] terrain 1024^2 + 512 spheres + 64 lights")
    scene = N.Scene.terrain(1, 1024, 512, 64)
    return scene, oracle_scene(scene)


def test_c3_batched_views_match_oracle(terrain):
    scene, orc = terrain
    frames = 2
    ctx = N.Context(scene, 8, 8)
    cams = [N.Camera.orbit(2 + v, 8, 64) for v in range(len(C3_SIZES))]
    views = [[c.view(f) for f in range(frames)] for c in cams]
    params = N.default_params(**PARAMS)
    ubos = [[N.make_ubo(views[v][f], views[v][max(f - 1, 0)], 45.0, w, h, bounces=4, frame=f + 1)
             for v, (w, h) in enumerate(C3_SIZES)] for f in range(frames)]
    batch = N.Batch(ctx, C3_SIZES)
    ms = batch.autotune(ubos[0], params, frames=2)
    say(f"[configs] C3 autotune -> config {batch.config} {N.Batch.config_info(batch.config)}, "
        f"{min(ms):.2f} ms/frame")
    rng = np.random.default_rng(17)
    for f in range(frames):
        frame = f + 1
        batch.render(ubos[f], params)
        ctx.sync()
        for v, (w, h) in enumerate(C3_SIZES):
            px = np.sort(rng.choice(w * h, size=min(2000, w * h), replace=False)).astype(np.uint32)
            say(f"[configs] C3 frame {frame} view {v} ({w}x{h})")
            valid = check_view_frame(orc, lambda which: batch.read(v, which), w, h, ubos[f][v], frame, px, params)
            assert valid > 0.3
    batch.close()
    ctx.close()
