"""Context isolation and resizing on the GPU: several xr_ctx in one process
keep their own frame buffers, history, acceleration structures and networks,
and a resized context renders exactly what a fresh context of that size
renders (InitReSTIRGITracer re-creates the frame resources for a new extent,
ApplicationReSTIRGI.cpp:304-374)."""
from __future__ import annotations

import numpy as np
import pytest

from xvk_ray_amd import native as N

from helpers import CORNELL_EYE, CORNELL_TARGET, cornell_scene, look_view, say, small_terrain

pytestmark = pytest.mark.gpu

BUFS = (N.XR_BUF_SAMPLES, N.XR_BUF_TEMPORAL, N.XR_BUF_SPATIAL, N.XR_BUF_OUTPUT)


def _restir_params():
    return N.default_params(temporal_m_cap=20, spatial_samples=5, spatial_radius=30.0, visibility=N.XR_VIS_TRACE)


def _cornell_ubos(W, H, frames=3):
    eyes = [(CORNELL_EYE[0] + 0.01 * k, CORNELL_EYE[1], CORNELL_EYE[2]) for k in range(frames)]
    views = [look_view(e, CORNELL_TARGET) for e in eyes]
    return [N.make_ubo(v, views[k - 1] if k else v, 60.0, W, H, bounces=4, frame=k + 1) for k, v in enumerate(views)]


def _terrain_ubos(W, H, frames=3):
    views = [look_view((0.2 + 0.01 * k, 0.4, 0.2), (0.5, 0.0, 0.5)) for k in range(frames)]
    return [N.make_ubo(v, views[k - 1] if k else v, 45.0, W, H, bounces=4, sky=1, frame=k + 1)
            for k, v in enumerate(views)]


def _plain_ubos(W, H, frames=2):
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    return [N.make_ubo(view, view, 60.0, W, H, spp=1, total=f, bounces=6, reuse=0, frame=f) for f in range(1, frames + 1)]


def _snapshot(ctx, bufs=BUFS):
    ctx.sync()
    return {b: ctx.read(b).tobytes() for b in bufs}


def test_two_contexts_interleaved_equal_separate_runs():
    """ReSTIR GI frames of two contexts (Cornell box with the LBVH, terrain with
    the GPU SAH tree) rendered alternately equal each context rendered alone."""
    Wa, Ha, Wb, Hb = 48, 32, 40, 24
    ca, cb = cornell_scene(), small_terrain()
    ua, ub = _cornell_ubos(Wa, Ha), _terrain_ubos(Wb, Hb)
    p = _restir_params()
    say(f"two contexts {Wa}x{Ha} cornell / {Wb}x{Hb} terrain, interleaved")
    a = N.Context(ca, Wa, Ha)
    b = N.Context(cb, Wb, Hb, builder=N.XR_BUILDER_SAH)
    for k in range(len(ua)):
        a.render(ua[k], N.XR_MODE_RESTIR, p)
        b.render(ub[k], N.XR_MODE_RESTIR, p)
    inter_a, inter_b = _snapshot(a), _snapshot(b)
    a.close()
    b.close()
    say("the same contexts alone")
    a1 = N.Context(ca, Wa, Ha)
    for u in ua:
        a1.render(u, N.XR_MODE_RESTIR, p)
    alone_a = _snapshot(a1)
    a1.close()
    b1 = N.Context(cb, Wb, Hb, builder=N.XR_BUILDER_SAH)
    for u in ub:
        b1.render(u, N.XR_MODE_RESTIR, p)
    alone_b = _snapshot(b1)
    b1.close()
    for which in BUFS:
        assert inter_a[which] == alone_a[which], f"cornell buffer {which} differs when interleaved"
        assert inter_b[which] == alone_b[which], f"terrain buffer {which} differs when interleaved"
    # the two scenes render different images (the comparison is not vacuous)
    assert inter_a[N.XR_BUF_OUTPUT] != inter_b[N.XR_BUF_OUTPUT][: len(inter_a[N.XR_BUF_OUTPUT])]


def _render_alone(scene, W, H, ubos, builder, p):
    ctx = N.Context(scene, W, H, builder=builder)
    for u in ubos:
        ctx.render(u, N.XR_MODE_RESTIR, p)
    snap = _snapshot(ctx)
    ctx.close()
    return snap


def test_contexts_on_two_host_threads_equal_separate_runs():
    """Two host threads, each with its own context and stream, render at the
    same time (the library's calls release the interpreter lock): each thread's
    frames equal the same context rendered alone."""
    import threading

    p = _restir_params()
    jobs = [
        (cornell_scene(), 96, 64, _cornell_ubos(96, 64, frames=4), N.XR_BUILDER_LBVH),
        (small_terrain(), 80, 48, _terrain_ubos(80, 48, frames=4), N.XR_BUILDER_SAH),
    ]
    want = [_render_alone(sc, W, H, u, b, p) for sc, W, H, u, b in jobs]
    for rep in range(3):
        got = [None, None]
        errors = []
        start = threading.Barrier(2)

        def work(k):
            try:
                sc, W, H, u, b = jobs[k]
                ctx = N.Context(sc, W, H, builder=b)
                start.wait()
                for x in u:
                    ctx.render(x, N.XR_MODE_RESTIR, p)
                got[k] = _snapshot(ctx)
                ctx.close()
            except Exception as e:  # reported below, in the main thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(k,), daemon=True) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in threads), "a render thread did not finish"
        assert not errors, errors
        say(f"two threads, repetition {rep}: done")
        for k in range(2):
            for which in BUFS:
                assert got[k][which] == want[k][which], f"job {k} buffer {which} differs on two threads"


def test_networks_are_per_context():
    """Each context holds its own radiance-cache network: different seeds give
    different predictions, a reset with the first seed restores them exactly,
    and training one context leaves the other's predictions unchanged."""
    scene = cornell_scene()
    a = N.Context(scene, 8, 8)
    b = N.Context(scene, 8, 8)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(256, N.XR_MLP_INPUTS)).astype(np.float32)
    y = rng.uniform(0, 1, size=(256, 3)).astype(np.float32)
    a.mlp_reset(1)
    b.mlp_reset(2)
    ya, yb = a.mlp_eval(x), b.mlp_eval(x)
    assert not np.array_equal(ya, yb)
    b.mlp_reset(1)
    assert np.array_equal(b.mlp_eval(x), ya)
    for _ in range(5):
        b.mlp_train_rows(x, y, lr=1e-2)
    assert not np.array_equal(b.mlp_eval(x), ya)
    assert np.array_equal(a.mlp_eval(x), ya), "training context b changed context a's network"
    a.close()
    b.close()


@pytest.mark.parametrize("sizes", [((48, 32), (33, 9), (64, 48)), ((17, 5), (80, 40), (1, 1))])
def test_resized_context_renders_like_a_fresh_one(sizes):
    """A context resized (down, then up) renders, from an empty history, the
    same ReSTIR GI and plain frames as a context created at that size."""
    scene = cornell_scene()
    p = _restir_params()
    pp = N.default_params(plain_max_depth=0)
    ctx = N.Context(scene, *sizes[0])
    for u in _cornell_ubos(*sizes[0], frames=2):
        ctx.render(u, N.XR_MODE_RESTIR, p)
    ctx.sync()
    for W, H in sizes[1:]:
        say(f"resize {ctx.width}x{ctx.height} -> {W}x{H}")
        ctx.resize(W, H)
        fresh = N.Context(scene, W, H)
        for u in _cornell_ubos(W, H):
            ctx.render(u, N.XR_MODE_RESTIR, p)
            fresh.render(u, N.XR_MODE_RESTIR, p)
        got, want = _snapshot(ctx), _snapshot(fresh)
        for which in BUFS:
            assert got[which] == want[which], f"buffer {which} differs after resize to {W}x{H}"
        for u in _plain_ubos(W, H):
            ctx.render(u, N.XR_MODE_PLAIN, pp)
            fresh.render(u, N.XR_MODE_PLAIN, pp)
        got = _snapshot(ctx, (N.XR_BUF_ACCUMULATION, N.XR_BUF_OUTPUT))
        want = _snapshot(fresh, (N.XR_BUF_ACCUMULATION, N.XR_BUF_OUTPUT))
        assert got == want, f"plain frames differ after resize to {W}x{H}"
        fresh.close()
    ctx.close()


def _device_free_bytes() -> int:
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return int(free.value)


def test_contexts_and_batches_release_device_memory():
    """Twelve create / build / render / destroy cycles (ReSTIR GI and plain
    frames, a GPU SAH rebuild, a cache frame, traced rays, a two-view batch
    replayed as a graph) leave the device's free memory where the third cycle
    left it: nothing the library allocates outlives its context."""
    scene = cornell_scene()
    W, H = 512, 384
    p = _restir_params()
    pc = _restir_params()
    pc.flags = N.XR_FLAG_MLP_CACHE
    pc.spatial_samples = 0
    rays = np.zeros(4096, N.RAY_DTYPE)
    rays["origin"] = (0.5, 0.5, -1.0)
    rays["direction"] = (0.0, 0.0, 1.0)
    rays["tmin"], rays["tmax"] = 1e-4, 1e4

    def cycle():
        ctx = N.Context(scene, W, H)
        ctx.build_accel(N.XR_BUILDER_SAH)
        for u in _cornell_ubos(W, H, frames=2):
            ctx.render(u, N.XR_MODE_RESTIR, p)
        ctx.mlp_reset(1)
        ctx.render(_cornell_ubos(W, H, frames=1)[0], N.XR_MODE_RESTIR, pc)
        ctx.render(_plain_ubos(W, H, frames=1)[0], N.XR_MODE_PLAIN, N.default_params())
        ctx.trace(rays)
        b = N.Batch(ctx, [(64, 48), (40, 40)])
        b.set_graph(True)
        for _ in range(2):
            b.render([_cornell_ubos(64, 48, frames=1)[0], _cornell_ubos(40, 40, frames=1)[0]], p)
        ctx.sync()
        live = _device_free_bytes()
        b.close()
        ctx.close()
        return live

    free = []
    for k in range(12):
        live = cycle()
        free.append(_device_free_bytes())
        say(f"cycle {k}: {free[-1] / 2**20:.1f} MiB free, {(free[-1] - live) / 2**20:.1f} MiB held while alive")
        # the measure sees the context's frame buffers (the 96-byte samples and
        # four reservoir slots alone are 90 MiB at 512x384)
        assert free[-1] - live >= 64 << 20
    drift = free[2] - free[-1]
    assert abs(drift) <= 2 << 20, f"device free memory moved by {drift / 2**20:.2f} MiB over nine cycles"


def test_injected_failures_leave_no_trace(monkeypatch):
    """Failures forced from the host (XR_FAULT_INJECT, read on every call): a
    context whose device allocation fails is refused with an out-of-memory
    error and gives back everything it had allocated; a frame whose launch
    fails is refused before any work is queued, so the context's later
    frames equal those of a context that never saw the failure."""
    scene = cornell_scene()
    W, H = 64, 48
    p = _restir_params()
    ubos = _cornell_ubos(W, H, frames=3)
    N.Context(scene, W, H).close()  # the runtime's one-time set-up happens outside the measure
    free0 = _device_free_bytes()
    monkeypatch.setenv("XR_FAULT_INJECT", "alloc")
    with pytest.raises(N.XrError) as e:
        N.Context(scene, W, H)
    assert e.value.code == N.XR_ERR_OUT_OF_MEMORY
    monkeypatch.delenv("XR_FAULT_INJECT")
    assert _device_free_bytes() == free0, "a refused context kept device memory"
    ctx = N.Context(scene, W, H)
    ctx.render(ubos[0], N.XR_MODE_RESTIR, p)
    monkeypatch.setenv("XR_FAULT_INJECT", "launch")
    with pytest.raises(N.XrError) as e:
        ctx.render(ubos[1], N.XR_MODE_RESTIR, p)
    assert e.value.code == N.XR_ERR_DEVICE and "injected" in str(e.value)
    monkeypatch.delenv("XR_FAULT_INJECT")
    for u in ubos[1:]:
        ctx.render(u, N.XR_MODE_RESTIR, p)
    got = _snapshot(ctx)
    ctx.close()
    want = _render_alone(scene, W, H, ubos, N.XR_BUILDER_LBVH, p)
    for which in BUFS:
        assert got[which] == want[which], f"buffer {which} differs after a refused frame"
