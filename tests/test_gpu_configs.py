"""The larger configurations against the oracle at their own workloads, on
pixel subsets (needs an MI355X).

  * 12 batched views of uneven size (4x 256x256, 4x 960x540, 2x 1920x1080,
    2x 1280x720) of the 2.75 M-triangle terrain, each on its own seeded
    camera path, one launch per pass for all views, replayed as a HIP graph
    with the autotuner's kernel configuration, and again with the BF16
    radiance cache (each view choosing its own training samples, one training
    step on all of them);
  * 1920x1080 ReSTIR GI with the BF16 radiance cache (65,536 training
    samples per frame) on the same terrain;
  * 3840x2160 ReSTIR GI at depth 6 with FP8 cache weights on the
    13.6 M-triangle terrain.

After every frame each view's buffers are read back and the oracle restates,
for a random subset of that view's pixels, the initial sample (from the
UBO), the temporal pass (from the frame's samples, motion, G-buffer and the
previous frame's reservoirs, which the frame leaves in place) and the spatial
pass (from the frame's samples and temporal reservoirs); the results must
agree bit for bit.  In cache frames the cached radiance of query and
training pixels is compared with the oracle's network on the weights after
the frame's training step instead (1e-2 of the output scale), and the
training choice (which pixels, how many) bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
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


def test_c3_batched_views_with_cache_match_oracle(terrain):
    """The 12 views of the batched workload with the BF16 radiance cache: each
    view chooses its own 65,536 training samples (or all its candidates) as a
    single-view frame would, one training step covers all views, and each
    view's subset matches the oracle (cache kinds and traced fields bit-exact,
    cached radiance from the updated network, temporal and spatial passes)."""
    scene, orc = terrain
    frames = 2
    ctx = N.Context(scene, 8, 8)
    cams = [N.Camera.orbit(2 + v, 8, 64) for v in range(len(C3_SIZES))]
    views = [[c.view(f) for f in range(frames)] for c in cams]
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE, **PARAMS)
    ubos = [[N.make_ubo(views[v][f], views[v][max(f - 1, 0)], 45.0, w, h, bounces=4, frame=f + 1)
             for v, (w, h) in enumerate(C3_SIZES)] for f in range(frames)]
    batch = N.Batch(ctx, C3_SIZES)
    rng = np.random.default_rng(29)
    for f in range(frames):
        frame = f + 1
        batch.reset_counters()
        batch.render(ubos[f], params)
        ctx.sync()
        w_after = ctx.mlp_params()
        total = 0
        for v, (w, h) in enumerate(C3_SIZES):
            px = np.sort(rng.choice(w * h, size=min(2000, w * h), replace=False)).astype(np.uint32)
            say(f"[configs] C3 + cache frame {frame} view {v} ({w}x{h})")
            read = lambda which, v=v: batch.read(v, which)  # noqa: E731
            nt, _ = check_cache_frame(read, w_after, orc, scene, w, h, ubos[f][v], px, params, fp8=False)
            assert 0 < nt <= params.mlp_train_pixels
            if w * h >= 1280 * 720:  # more candidates than the target
                assert nt == params.mlp_train_pixels
            _passes_after(read, orc, w, h, frame, px, params)
            total += nt
        say(f"[configs] C3 + cache frame {frame}: {total} training samples over {len(C3_SIZES)} views")
        assert batch.counters()[4] == total
    batch.close()
    ctx.close()


@pytest.mark.parametrize("cache", [False, True])
def test_c3_long_history(terrain, cache):
    """The batched workload over 64 frames of each view's camera path, replayed
    as graphs under the autotuned configuration: every view's reservoirs stay
    finite with M within the caps and its image opaque; without the cache the
    64th frame of every view still equals the oracle on a pixel subset, with
    the cache the shared network stays finite and each step trains on the
    union of the views' samples (at most 65,536 per view, all of them for the
    views with more candidates than that)."""
    scene, orc = terrain
    frames = 64
    ctx = N.Context(scene, 8, 8)
    cams = [N.Camera.orbit(2 + v, 8, frames) for v in range(len(C3_SIZES))]
    views = [[c.view(f) for f in range(frames)] for c in cams]
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE if cache else 0, **PARAMS)
    ubos = [[N.make_ubo(views[v][f], views[v][max(f - 1, 0)], 45.0, w, h, bounces=4, frame=f + 1)
             for v, (w, h) in enumerate(C3_SIZES)] for f in range(frames)]
    batch = N.Batch(ctx, C3_SIZES)
    batch.set_graph(True)
    if not cache:
        batch.autotune(ubos[0], params, frames=2)
        batch.reset()
    else:
        ctx.mlp_reset(1)
    w0 = ctx.mlp_params().copy() if cache else None
    batch.reset_counters()
    for f in range(frames):
        batch.render(ubos[f], params)
    ctx.sync()
    say(f"[configs] C3 {'+ cache ' if cache else ''}after {frames} frames (config {batch.config})")
    rng = np.random.default_rng(31)
    for v, (w, h) in enumerate(C3_SIZES):
        for which, cap in ((N.XR_BUF_TEMPORAL, params.temporal_m_cap), (N.XR_BUF_SPATIAL, params.spatial_m_cap)):
            r = batch.read(v, which)
            for fld in ("w_sum", "w"):
                assert np.isfinite(r[fld]).all() and (r[fld] >= 0).all(), f"view {v} buffer {which} {fld}"
            assert (r["m"] <= cap).all(), f"view {v} buffer {which} M above {cap}"
        assert ((batch.read(v, N.XR_BUF_OUTPUT) >> 24) == 0xFF).all(), f"view {v} output alpha"
        if not cache:
            px = np.sort(rng.choice(w * h, size=min(2000, w * h), replace=False)).astype(np.uint32)
            valid = check_view_frame(orc, lambda which, v=v: batch.read(v, which), w, h, ubos[-1][v], frames, px,
                                     params)
            assert valid > 0.3
    if cache:
        wl = ctx.mlp_params()
        assert np.isfinite(wl).all() and not np.array_equal(wl, w0)
        trained = batch.counters()[4]
        big = sum(1 for w, h in C3_SIZES if w * h >= 1280 * 720)
        say(f"[configs] C3 + cache: {trained / frames:.0f} training samples per step")
        assert frames * big * params.mlp_train_pixels <= trained <= frames * len(C3_SIZES) * params.mlp_train_pixels
    batch.close()
    ctx.close()


def test_c5_long_run_stays_finite():
    """The 4K workload's whole 128-frame camera path (13.6 M triangles, depth 6,
    FP8 radiance cache with the five spatial neighbours): after the last frame
    the reservoirs keep their invariants, the image is opaque, the radiance
    and the network are finite, and every frame trained on 65,536 samples."""
    say("[configs] C5 scene (long run)")
    scene = N.Scene.terrain(1, 2048, 4096, 256)
    W, H = 3840, 2160
    ctx = N.Context(scene, W, H, builder=N.XR_BUILDER_SAH)
    cam = N.Camera.orbit(2, 8, 128)
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE | N.XR_FLAG_MLP_FP8 | N.XR_FLAG_RADIANCE, **PARAMS)
    ctx.mlp_reset(1)
    ctx.reset_counters()
    frames = 128
    prev = cam.view(0)
    for k in range(frames):
        view = cam.view(k)
        ctx.render(N.make_ubo(view, prev, 45.0, W, H, bounces=6, frame=k + 1), N.XR_MODE_RESTIR, params)
        prev = view
    ctx.sync()
    trained = ctx.counters_all()[4]
    say(f"[configs] C5 after {frames} frames: {trained / frames:.0f} training samples per frame")
    assert trained == frames * params.mlp_train_pixels
    assert np.isfinite(ctx.mlp_params()).all()
    assert np.isfinite(ctx.read(N.XR_BUF_RADIANCE)).all()
    for which, cap in ((N.XR_BUF_TEMPORAL, params.temporal_m_cap), (N.XR_BUF_SPATIAL, params.spatial_m_cap)):
        r = ctx.read(which)
        for fld in ("w_sum", "w"):
            assert np.isfinite(r[fld]).all() and (r[fld] >= 0).all(), f"buffer {which} {fld}"
        assert (r["m"] <= cap).all(), f"buffer {which} M above {cap}"
        del r
    assert ((ctx.read(N.XR_BUF_OUTPUT) >> 24) == 0xFF).all()
    ctx.close()


def _features(samples, mats, lo, inv, idx):
    from tests.test_gpu_mlp import _features as f
    return f(samples, mats, lo, inv, idx)


def check_cache_frame(read, w_after, orc, scene, W, H, ubo, px, params, fp8):
    """One radiance-cache frame of one view on the pixels px (`read(which)`
    returns the view's buffer): every traced field and the cache kind
    bit-exact; query and training pixels carry the updated network's
    prediction (within 1e-2 of the output scale).  Returns the view's whole
    training-set size as the oracle chooses it and the cached pixels in px."""
    samples = read(N.XR_BUF_SAMPLES).copy()
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, pixels=px, cache=(params.mlp_train_pixels,
                                                                           params.mlp_spread))
    rs = ref["samples"]
    train_all = ref["train"].astype(bool)
    g, r = samples[px].copy(), rs[px].copy()
    kind = r["pad2"]
    assert np.array_equal(g["pad2"], kind)
    cached = (kind == 1) | (kind == 2)
    g["l_out"][cached] = 0
    r["l_out"][cached] = 0
    _eq("cache-frame samples", g, r)
    # the network's radiance at query / training pixels (a bounded sample of them)
    verts, _, mats, _ = scene.arrays()
    lo = verts["pos"].min(0).astype(np.float32)
    inv = (np.float32(1.0) / (verts["pos"].max(0) - lo).astype(np.float32)).astype(np.float32)
    idx = px[cached][:3000]
    pred = np.maximum(B.mlp_forward(w_after, _features(samples, mats, lo, inv, idx), mode=2 if fp8 else 1), 0.0)
    got = samples["l_out"][idx].astype(np.float64)
    scale = max(1.0, np.abs(pred).max())
    assert np.abs(got - pred).max() <= 1e-2 * scale, float(np.abs(got - pred).max())
    return int(train_all.sum()), int(cached.sum())


def _passes_after(read, orc, W, H, frame, px, params):
    from oracle.binding import temporal as orc_temporal
    n = W * H
    samples = read(N.XR_BUF_SAMPLES).copy()
    # the samples' spare words carry the cache kind and x_s material; a
    # reservoir has no room for them (its planes' spare words hold w_sum, W, M)
    samples["pad1"] = 0
    samples["pad2"] = 0
    motion = read(N.XR_BUF_MOTION).copy()
    gpos = np.ascontiguousarray(read(N.XR_BUF_GBUF_POS))
    t_gpu = read(N.XR_BUF_TEMPORAL).copy()
    s_gpu = read(N.XR_BUF_SPATIAL).copy()
    cur, wr = frame % 2, 1 - frame % 2
    t_ref, s_ref = t_gpu.copy(), s_gpu.copy()
    orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, t_ref, s_ref, np.zeros(n, np.uint32),
                 pixels=px)
    _eq("temporal reservoirs", t_gpu[px + cur * n], t_ref[px + cur * n])
    s_ref = s_gpu.copy()
    orc.spatial(W, H, frame, 1, params.spatial_samples, params.spatial_radius, params.spatial_m_cap,
                params.visibility, samples, t_gpu, s_ref, np.zeros(n, np.uint32), pixels=px)
    _eq("spatial reservoirs", s_gpu[px + wr * n], s_ref[px + wr * n])


def test_c4_cache_1080p_matches_oracle(terrain):
    """1920x1080 ReSTIR GI with the BF16 radiance cache on the 2.75 M-triangle
    terrain, 65,536 training samples per frame."""
    scene, orc = terrain
    W, H = 1920, 1080
    ctx = N.Context(scene, W, H)
    cam = N.Camera.orbit(2, 8, 128)
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE, **PARAMS)
    assert params.mlp_train_pixels == 65536
    rng = np.random.default_rng(19)
    px = np.sort(rng.choice(W * H, size=20000, replace=False)).astype(np.uint32)
    prev = cam.view(0)
    for k in range(4):
        view = cam.view(k)
        ubo = N.make_ubo(view, prev, 45.0, W, H, bounces=4, frame=k + 1)
        prev = view
        ctx.reset_counters()
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        ctx.sync()
        if k >= 2:
            say(f"[configs] C4 frame {k + 1}")
            nt, nc = check_cache_frame(ctx.read, ctx.mlp_params(), orc, scene, W, H, ubo, px, params, fp8=False)
            assert ctx.counters_all()[4] == nt  # the frame's whole training set, as the oracle chose it
            say(f"[configs] C4 frame {k + 1}: {nt} training samples, {nc} cached pixels in the subset")
            assert nt == 65536
            _passes_after(ctx.read, orc, W, H, k + 1, px, params)


def test_c5_4k_fp8_depth6_matches_oracle():
    """3840x2160 ReSTIR GI, max depth 6, FP8 radiance cache, on the
    13.6 M-triangle terrain (2048^2 heightfield + 4,096 icospheres + 256 lights)."""
    say("[configs] C5 scene")
    scene = N.Scene.terrain(1, 2048, 4096, 256)
    orc = oracle_scene(scene)
    W, H = 3840, 2160
    ctx = N.Context(scene, W, H)
    cam = N.Camera.orbit(2, 8, 128)
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE | N.XR_FLAG_MLP_FP8, **PARAMS)
    rng = np.random.default_rng(23)
    px = np.sort(rng.choice(W * H, size=20000, replace=False)).astype(np.uint32)
    prev = cam.view(0)
    for k in range(2):
        view = cam.view(k)
        ubo = N.make_ubo(view, prev, 45.0, W, H, bounces=6, frame=k + 1)
        prev = view
        ctx.reset_counters()
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        ctx.sync()
        say(f"[configs] C5 frame {k + 1}")
        nt, nc = check_cache_frame(ctx.read, ctx.mlp_params(), orc, scene, W, H, ubo, px, params, fp8=True)
        assert ctx.counters_all()[4] == nt
        say(f"[configs] C5 frame {k + 1}: {nt} training samples, {nc} cached pixels in the subset")
        assert nt == 65536
        _passes_after(ctx.read, orc, W, H, k + 1, px, params)
