"""The headless frame driver's bookkeeping (PathTracer::DrawFrame,
src/pathTracer.cpp:71-79) through xrh_app_draw_frame (needs an MI355X).

  * with accumulation on, spp per frame is clamp(maxSpp - totalSamples, 0, spp):
    it falls to the remainder and then to 0 at maxSpp;
  * a camera move (rebuildRays, :88) restarts accumulation on the next frame;
  * with accumulation off every frame restarts (totalSamples = spp);
  * the frame counter wraps at 10,000 (:79).
The UBO of every frame is read back, and the GPU's accumulation image over
the sequence equals the oracle's plain path tracer run on the same UBOs.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import CORNELL_EYE, CORNELL_TARGET, bytes_equal_rows, describe_mismatch, oracle_scene, say
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _run(app, orc, W, H, frames):
    ctx = app.context()
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    seen = []
    for k in range(frames):
        app.draw_frame()
        ctx.sync()
        ubo = app.last_ubo()
        seen.append((ubo.spp, ubo.total_number_of_samples, ubo.current_frame))
        orc.plain(N.ubo_to_bytes(ubo), W, H, 0, accum, out)
        ga = ctx.read(N.XR_BUF_ACCUMULATION)
        eq = bytes_equal_rows(ga, accum)
        assert eq.all(), describe_mismatch(f"accumulation, frame {k + 1}", ga, accum, eq)
    return seen


def test_accumulation_clamps_at_max_spp_and_restarts_after_a_move():
    scene = N.Scene.cornell(1)
    W, H = 40, 32
    moved = (CORNELL_EYE[0] + 0.05, CORNELL_EYE[1], CORNELL_EYE[2])
    eyes = [CORNELL_EYE] * 4 + [moved] * 4  # path frame 4 moves the camera
    cam = N.Camera.keyframes(eyes, [CORNELL_TARGET] * 8)
    app = N.App(scene, cam, W, H, mode=N.XR_MODE_PLAIN, spp=2, max_spp=5, fov=45.0, accumulate=True)
    say("[driver] accumulation 2 spp per frame up to 5, camera moved at path frame 4")
    seen = _run(app, oracle_scene(scene), W, H, 8)
    # frames 1-3: 2 + 2 + 1 samples; frame 4: none left; frame 5 renders the
    # moved view with none (the move is seen after the bookkeeping); frame 6
    # restarts from 0
    assert [(s, t) for s, t, _ in seen] == [(2, 2), (2, 4), (1, 5), (0, 5), (0, 5), (2, 2), (2, 4), (1, 5)]
    assert [f for _, _, f in seen] == list(range(1, 9))
    app.close()


def test_without_accumulation_every_frame_restarts_and_counter_wraps():
    scene = N.Scene.cornell(1)
    W, H = 8, 8
    cam = N.Camera.static(CORNELL_EYE, CORNELL_TARGET)
    app = N.App(scene, cam, W, H, mode=N.XR_MODE_PLAIN, spp=3, max_spp=1000, fov=45.0, accumulate=False)
    say("[driver] no accumulation, 3 spp")
    seen = _run(app, oracle_scene(scene), W, H, 4)
    assert [(s, t) for s, t, _ in seen] == [(3, 3)] * 4
    for _ in range(10000 - 4 - 1):
        app.draw_frame()
    app.context().sync()
    assert app.last_ubo().current_frame == 9999
    app.draw_frame()
    assert app.last_ubo().current_frame == 0  # (currentFrameTotal + 1) % 10000
    app.context().sync()
    app.close()
