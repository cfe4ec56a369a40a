"""Batched views through the C ABI (needs an MI355X).

A batch renders several cameras on one scene with one launch per pass and a
replayed HIP graph.  Each view must equal the same view rendered alone by a
single-view context (bit for bit: same kernels' arithmetic), for every kernel
configuration the autotuner chooses from, with and without the graph, and
across graph replays separated by host reads.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import look_view, say, small_terrain
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (40, 40), (96, 32)]
EYES = [((0.2, 0.45, 0.1), (0.6, 0.1, 0.7)), ((0.8, 0.4, 0.2), (0.3, 0.05, 0.6)), ((0.5, 0.5, 0.9), (0.5, 0.1, 0.2))]
PARAMS = dict(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)
BUFFERS = [N.XR_BUF_SAMPLES, N.XR_BUF_TEMPORAL, N.XR_BUF_SPATIAL, N.XR_BUF_OUTPUT, N.XR_BUF_MOTION]


def ubo_sequence(frames: int):
    """UBOs per frame per view: each camera drifts a little every frame."""
    seq = []
    prev = [None] * len(SIZES)
    for f in range(1, frames + 1):
        row = []
        for v, ((eye, tgt), (w, h)) in enumerate(zip(EYES, SIZES)):
            e = np.array(eye) + 0.004 * f * np.array([1.0, 0.0, 0.5])
            view = look_view(tuple(e), tgt)
            row.append(N.make_ubo(view, prev[v] if prev[v] is not None else view, 45.0, w, h, bounces=4, frame=f))
            prev[v] = view
        seq.append(row)
    return seq


@pytest.fixture(scope="module")
def scene():
    return small_terrain()


@pytest.fixture(scope="module")
def singles(scene):
    """Every view rendered alone, frame by frame: {frame: [per-view buffer dict]}."""
    seq = ubo_sequence(3)
    ctxs = [N.Context(scene, w, h) for w, h in SIZES]
    p = N.default_params(**PARAMS)
    out = {}
    for f, row in enumerate(seq, 1):
        out[f] = []
        for ctx, ubo in zip(ctxs, row):
            ctx.render(ubo, N.XR_MODE_RESTIR, p)
            ctx.sync()
            out[f].append({b: ctx.read(b).copy() for b in BUFFERS})
    return seq, out


def _check(batch, ref, label):
    for v in range(len(SIZES)):
        for b in BUFFERS:
            got = batch.read(v, b)
            assert got.tobytes() == ref[v][b].tobytes(), f"{label}: view {v} buffer {b}"


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("cfg", range(18))
def test_batch_equals_single_views(scene, singles, cfg, graph):
    seq, ref = singles
    assert N.Batch.config_count() == 18
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, SIZES)
    batch.set_config(cfg)
    batch.set_graph(graph)
    say(f"[batch] cfg {cfg} {N.Batch.config_info(cfg)} persistent {N.Batch.config_persistent(cfg)} graph {graph}")
    p = N.default_params(**PARAMS)
    for f, row in enumerate(seq, 1):
        batch.render(row, p)
        ctx.sync()
        _check(batch, ref[f], f"cfg {cfg} frame {f}")
    ext, shadow, valid = batch.counters()[:3]
    assert ext >= valid > 0 and shadow > 0
    batch.close()
    ctx.close()


def test_graph_replays_with_host_work_between(scene, singles):
    """Many replays of one captured graph, with reads and a parameter change
    (which captures a new graph) in between, against eager launches."""
    seq = ubo_sequence(6)
    ctx = N.Context(scene, 8, 8)
    g = N.Batch(ctx, SIZES)
    e = N.Batch(ctx, SIZES)
    e.set_graph(False)
    for f, row in enumerate(seq, 1):
        p = N.default_params(**dict(PARAMS, spatial_samples=5 if f <= 3 else 3))
        g.render(row, p)
        e.render(row, p)
        ctx.sync()
        for v in range(len(SIZES)):
            for b in BUFFERS:
                assert g.read(v, b).tobytes() == e.read(v, b).tobytes(), (f, v, b)
    assert g.frame_time() > 0.0


def test_autotune_picks_fastest_and_clears(scene):
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, SIZES)
    row = ubo_sequence(1)[0]
    ms = batch.autotune(row, N.default_params(**PARAMS), frames=2)
    say(f"[batch] autotune ms {['%.3f' % x for x in ms]} -> {batch.config}")
    assert len(ms) == 18 and all(x > 0 for x in ms)
    assert batch.config == int(np.argmin(ms))
    for v in range(len(SIZES)):
        assert not batch.read(v, N.XR_BUF_TEMPORAL, N.XR_LAYOUT_NATIVE).any()
    assert batch.counters() == (0, 0, 0, 0, 0)


def test_batch_rejects_bad_input(scene):
    ctx = N.Context(scene, 8, 8)
    with pytest.raises(N.XrError):
        N.Batch(ctx, [(0, 16)])
    batch = N.Batch(ctx, SIZES)
    p = N.default_params(**PARAMS)
    p.passes = 1  # batched views run every pass
    with pytest.raises(N.XrError):
        batch.render(ubo_sequence(1)[0], p)
    with pytest.raises(N.XrError):
        batch.set_config(99)
    with pytest.raises(N.XrError):
        batch.read(7, N.XR_BUF_OUTPUT)


def test_graph_recaptured_after_new_textures_and_rebuilt_tree():
    """A captured graph holds the scene's device pointers by value: new
    textures (reallocated) or a rebuilt tree must make the next frame capture
    again.  The graph batch must keep equal to an eager batch on the same
    context across both changes."""
    from tests.helpers import CORNELL_TARGET
    from tests.test_gpu_parity import textured_cornell
    scene = textured_cornell()
    ctx = N.Context(scene, 8, 8)
    sizes = [(48, 40), (32, 24)]
    eyes = [(0.5, 0.5, -1.2), (0.45, 0.55, -1.1)]
    g = N.Batch(ctx, sizes)
    e = N.Batch(ctx, sizes)
    e.set_graph(False)
    p = N.default_params(**PARAMS)
    rng = np.random.default_rng(8)
    for f in range(1, 7):
        if f == 3:  # new texture contents and sizes: the texture arrays are reallocated
            say("[batch] new textures")
            ctx.set_textures([(7, 5, rng.integers(0, 256, 7 * 5 * 4, dtype=np.uint8)),
                              (3, 3, rng.integers(0, 256, 3 * 3 * 4, dtype=np.uint8)),
                              (1, 1, np.array([10, 20, 30, 255], np.uint8))])
        if f == 5:  # a rebuilt tree: nodes and triangles are freed and allocated again
            say("[batch] rebuilt tree")
            ctx.build_accel(N.XR_BUILDER_PLOC)
        row = [N.make_ubo(look_view(eye, CORNELL_TARGET), look_view(eye, CORNELL_TARGET), 60.0, w, h, bounces=4,
                          frame=f) for eye, (w, h) in zip(eyes, sizes)]
        g.render(row, p)
        e.render(row, p)
        ctx.sync()
        for v in range(len(sizes)):
            for b in BUFFERS:
                assert g.read(v, b).tobytes() == e.read(v, b).tobytes(), (f, v, b)


TINY_SIZES = [(1, 1), (33, 9), (9, 1), (1, 17), (64, 48)]
TINY_EYES = EYES + [((0.3, 0.5, 0.4), (0.6, 0.1, 0.6)), ((0.7, 0.45, 0.8), (0.4, 0.1, 0.3))]


@pytest.mark.parametrize("cfg", [0, 3, 6, 9, 14])
def test_batch_views_smaller_than_a_tile(scene, cfg):
    """Views of one pixel, one row, one column and sizes that fill no tile, in
    grid and persistent launch forms with the graph: each view equals the same
    view rendered alone, over three frames."""
    frames = 3
    ctxs = [N.Context(scene, w, h) for w, h in TINY_SIZES]
    p = N.default_params(**PARAMS)
    ctx = N.Context(scene, 8, 8)
    batch = N.Batch(ctx, TINY_SIZES)
    batch.set_config(cfg)
    say(f"[batch] tiny views, cfg {cfg} {N.Batch.config_info(cfg)} persistent {N.Batch.config_persistent(cfg)}")
    prev = [None] * len(TINY_SIZES)
    for f in range(1, frames + 1):
        row = []
        for v, ((eye, tgt), (w, h)) in enumerate(zip(TINY_EYES, TINY_SIZES)):
            view = look_view(tuple(np.array(eye) + 0.004 * f), tgt)
            row.append(N.make_ubo(view, prev[v] if prev[v] is not None else view, 45.0, w, h, bounces=4, frame=f))
            prev[v] = view
        batch.render(row, p)
        for c, ubo in zip(ctxs, row):
            c.render(ubo, N.XR_MODE_RESTIR, p)
        ctx.sync()
        for v, c in enumerate(ctxs):
            c.sync()
            for b in BUFFERS:
                assert batch.read(v, b).tobytes() == c.read(b).tobytes(), f"cfg {cfg} frame {f} view {v} buffer {b}"
    batch.close()
    ctx.close()
