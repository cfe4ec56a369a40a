"""Batched views with the MLP radiance cache (needs an MI355X).

Each view of a batch makes its own training choice, exactly as a single-view
frame would (min(T, candidates) samples ranked in its pixel order, rotated by
its frame number); one Adam step then trains on the samples of all views
together (the mean relative-L2 loss over their union), and every view's query
and training pixels carry the updated network's radiance before its
temporal and spatial passes.

  * per view and frame: termination kinds, materials at x_s and every traced
    field bit-exact against the oracle, the cached radiance within 1e-2 of the
    oracle's network on the weights after the step, the temporal and spatial
    reservoirs bit-exact on the frame's samples;
  * the first step after a reset moves every weight by -lr * sign(g): its
    signs must be those of the oracle's gradient over the union of the views'
    training samples (a per-view normalisation or a missing view changes
    them on the smaller gradients, a wrong view's features scramble them);
  * the replayed graphs follow a changed training target.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
from tests.helpers import bytes_equal_rows, describe_mismatch, oracle_scene, say, small_terrain
from tests.test_gpu_batch import PARAMS, SIZES, ubo_sequence
from tests.test_gpu_mlp import _features
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu

SPREAD = 0.01
T = 300  # fewer than the candidates of the larger views, more than some


@pytest.fixture(scope="module")
def scene():
    return small_terrain()


def _eq(name, got, ref):
    eq = bytes_equal_rows(got, ref)
    assert eq.all(), describe_mismatch(name, got, ref, eq)


def _norm(scene):
    verts = scene.arrays()[0]
    lo = verts["pos"].min(0).astype(np.float32)
    inv = (np.float32(1.0) / (verts["pos"].max(0) - lo).astype(np.float32)).astype(np.float32)
    return lo, inv


def _params(**kw):
    return N.default_params(flags=N.XR_FLAG_MLP_CACHE, mlp_train_pixels=T, mlp_spread=SPREAD, **dict(PARAMS, **kw))


def _check_view(batch, orc, scene, v, ubo, frame, params, w_after, lo, inv):
    """One view after a cache frame; returns its training-sample count."""
    from oracle.binding import temporal as orc_temporal
    W, H = SIZES[v]
    n = W * H
    _, _, mats, _ = scene.arrays()
    gpu = batch.read(v, N.XR_BUF_SAMPLES).copy()
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, cache=(T, SPREAD))["samples"]
    kind = ref["pad2"]
    assert np.array_equal(gpu["pad2"], kind) and np.array_equal(gpu["pad1"], ref["pad1"]), f"view {v} kinds"
    query, train = kind == 1, kind == 2
    g2, r2 = gpu.copy(), ref.copy()
    g2["l_out"][query | train] = 0
    r2["l_out"][train] = 0
    _eq(f"view {v} cache-frame samples", g2, r2)
    idx = np.nonzero(query | train)[0]
    if len(idx):
        pred = np.maximum(B.mlp_forward(w_after, _features(ref, mats, lo, inv, idx), mode=1), 0.0)
        got = gpu["l_out"][idx].astype(np.float64)
        scale = max(1.0, np.abs(pred).max())
        assert np.abs(got - pred).max() <= 1e-2 * scale, (v, float(np.abs(got - pred).max()))
    # temporal and spatial passes on the frame's samples (the spare words carry
    # the cache kind and x_s material, which a reservoir has no room for)
    samples = gpu.copy()
    samples["pad1"] = 0
    samples["pad2"] = 0
    motion = batch.read(v, N.XR_BUF_MOTION).copy()
    gpos = np.ascontiguousarray(batch.read(v, N.XR_BUF_GBUF_POS))
    t_gpu = batch.read(v, N.XR_BUF_TEMPORAL).copy()
    s_gpu = batch.read(v, N.XR_BUF_SPATIAL).copy()
    cur, wr = frame % 2, 1 - frame % 2
    px = np.arange(n, dtype=np.uint32)
    t_ref, s_ref = t_gpu.copy(), s_gpu.copy()
    orc_temporal(W, H, frame, params.temporal_m_cap, samples, motion, gpos, t_ref, s_ref, np.zeros(n, np.uint32),
                 pixels=px)
    _eq(f"view {v} temporal reservoirs", t_gpu[px + cur * n], t_ref[px + cur * n])
    s_ref = s_gpu.copy()
    orc.spatial(W, H, frame, 1, params.spatial_samples, params.spatial_radius, params.spatial_m_cap,
                params.visibility, samples, t_gpu, s_ref, np.zeros(n, np.uint32), pixels=px)
    _eq(f"view {v} spatial reservoirs", s_gpu[px + wr * n], s_ref[px + wr * n])
    return int(train.sum())


@pytest.mark.parametrize("graph", [True, False])
def test_batch_cache_frames_match_oracle(scene, graph):
    orc = oracle_scene(scene)
    lo, inv = _norm(scene)
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, SIZES)
    batch.set_graph(graph)
    params = _params()
    seq = ubo_sequence(3)
    for f, row in enumerate(seq, 1):
        say(f"[batch-cache] graph {graph} frame {f}")
        batch.reset_counters()
        batch.render(row, params)
        ctx.sync()
        w_after = ctx.mlp_params()
        trained = [_check_view(batch, orc, scene, v, row[v], f, params, w_after, lo, inv) for v in range(len(SIZES))]
        counts = batch.counters()
        say(f"[batch-cache] frame {f}: training samples per view {trained}, queries {counts[3]}")
        assert counts[4] == sum(trained) and max(trained) == T and counts[3] > 0
    batch.close()
    ctx.close()


def test_batch_cache_step_trains_on_all_views(scene):
    orc = oracle_scene(scene)
    lo, inv = _norm(scene)
    _, _, mats, _ = scene.arrays()
    lr = 1e-3
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, SIZES)
    ctx.mlp_reset(5)
    w0 = ctx.mlp_params()
    w0[B.mlp_layer_offsets(5)[1]:] = 1.0  # output biases away from 0: no dead outputs on the first step
    ctx.mlp_params(w0)
    row = ubo_sequence(1)[0]
    say("[batch-cache] first cache frame of the batch after a reset")
    batch.render(row, _params(mlp_lr=lr))
    ctx.sync()
    w1 = ctx.mlp_params().astype(np.float64)
    xs, ys = [], []
    for v, (W, H) in enumerate(SIZES):
        ref = orc.initial_sample(N.ubo_to_bytes(row[v]), W, H, cache=(T, SPREAD))["samples"]
        train = np.nonzero(ref["pad2"] == 2)[0]
        xs.append(_features(ref, mats, lo, inv, train))
        ys.append(ref["l_out"][train].astype(np.float32))
    x, y = np.concatenate(xs), np.concatenate(ys)
    say(f"[batch-cache] {len(x)} training samples over {len(SIZES)} views")
    assert len(x) > 400
    _, g = B.mlp_loss_grad(w0, x, y, mode=1)
    step = w1 - w0.astype(np.float64)
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    huge = np.abs(g) > 1e-1 * np.abs(g).max()
    assert big.sum() > 500
    agree = np.sign(step[big]) == -np.sign(g[big])
    assert agree.mean() >= 0.995, agree.mean()
    assert np.all(np.sign(step[huge]) == -np.sign(g[huge]))
    assert np.allclose(np.abs(step[big]), lr, rtol=2e-2)
    batch.close()
    ctx.close()


def test_batch_cache_graph_follows_a_new_training_target(scene):
    """The captured initial-sample graph bakes in the training target: a
    changed target must capture it again (every view then trains the new
    number, as the oracle chooses them)."""
    orc = oracle_scene(scene)
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, SIZES)
    seq = ubo_sequence(3)
    for f, (row, t) in enumerate(zip(seq, (T, T, 100)), 1):
        batch.reset_counters()
        batch.render(row, N.default_params(flags=N.XR_FLAG_MLP_CACHE, mlp_train_pixels=t, mlp_spread=SPREAD,
                                           **PARAMS))
        ctx.sync()
        trained = 0
        for v, (W, H) in enumerate(SIZES):
            ref = orc.initial_sample(N.ubo_to_bytes(row[v]), W, H, cache=(t, SPREAD))["samples"]
            assert np.array_equal(batch.read(v, N.XR_BUF_SAMPLES)["pad2"], ref["pad2"]), (f, v)
            trained += int((ref["pad2"] == 2).sum())
        say(f"[batch-cache] frame {f}: target {t}, {trained} training samples")
        assert batch.counters()[4] == trained and trained == len(SIZES) * t
    batch.close()
    ctx.close()
