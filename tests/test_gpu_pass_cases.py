"""The closed-form pass inputs through the HIP passes (needs an MI355X).

The inputs of tests/test_cpu_oracle_closed_forms.py are written into a
context's buffers through xr_write_buffer and the pass runs on the device;
its outputs must equal the oracle's bit for bit and meet the same closed
forms: the emitter ignoring its texture, motion vectors of a translated
camera, the temporal M-cap rescale and W, the spatial Jacobian inside and at
both clamps, and the visibility ray leaving from the selected sample.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from tests.helpers import bytes_equal_rows, cornell_scene, describe_mismatch, look_view, oracle_scene, say
from tests.pass_cases import (JACOBIAN_CASES, floor_light_textured, jacobian, occluder_scene, spatial_layout,
                              translation_view, wall_scene)
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _initial(scene, ubo, W, H):
    ctx = N.Context(scene, W, H)
    p = N.default_params()
    p.passes = N.XR_PASS_INITIAL
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    ref = oracle_scene(scene).initial_sample(N.ubo_to_bytes(ubo), W, H)
    return ctx, ref


def test_emitter_ignores_texture_on_device():
    scene = floor_light_textured()
    W, H = 16, 12
    view = look_view((0, 0.5, 0), (0.0, 0.0, -1.0))
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    say("[cases] textured floor under a textured light")
    ctx, ref = _initial(scene, ubo, W, H)
    s = ctx.read(N.XR_BUF_SAMPLES)
    eq = bytes_equal_rows(s, ref["samples"])
    assert eq.all(), describe_mismatch("samples", s, ref["samples"], eq)
    floor = s["material_id"] == 0
    assert floor.sum() > W * H // 2 and np.all(s["l_out"][floor] == np.float32(8.0))


@pytest.mark.parametrize("delta", [(0.0, 0.0), (0.03, -0.02)])
def test_motion_translated_camera_on_device(delta):
    from tests.test_cpu_oracle_closed_forms import _jitter, _motion_f64
    from oracle.binding import Oracle
    W, H, fov, frame = 32, 24, 60.0, 3
    eye = np.array([delta[0], delta[1], 0.0])
    ubo = N.make_ubo(translation_view(eye), translation_view((0.0, 0.0, 0.0)), fov, W, H, bounces=1, frame=frame)
    say(f"[cases] motion, camera translated by {delta}")
    ctx, ref = _initial(wall_scene(), ubo, W, H)
    mv = ctx.read(N.XR_BUF_MOTION)
    assert np.array_equal(mv, ref["motion"])
    xi = _jitter(Oracle.lib(), W, H, frame)
    tan_h = math.tan(math.radians(fov) / 2)
    got = _motion_f64(mv)
    want_x = (0.5 - xi[:, 0]) / W - 0.5 * delta[0] / (tan_h * W / H)
    want_y = (0.5 - xi[:, 1]) / H + 0.5 * delta[1] / tan_h
    assert np.all(np.abs(got[:, 0] - want_x) <= np.abs(want_x) * 2.0 ** -11 + 2e-6)
    assert np.all(np.abs(got[:, 1] - want_y) <= np.abs(want_y) * 2.0 ** -11 + 2e-6)


def test_temporal_cap_on_device():
    from oracle.binding import temporal
    from tests.test_cpu_oracle_passes import _synthetic_frame
    W, H, cap, frame = 40, 30, 20, 3
    n = W * H
    s, t, gpos, motion, _ = _synthetic_frame(W, H, 2, invalid_frac=0.03)
    rng = np.random.default_rng(4)
    t["s"][:n] = s[rng.permutation(n)]
    t["m"][:n] = rng.integers(0, 40, n)
    ctx = N.Context(cornell_scene(), W, H)
    view = look_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5))
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=frame)
    for which, data in [(N.XR_BUF_SAMPLES, s), (N.XR_BUF_TEMPORAL, t), (N.XR_BUF_GBUF_POS, gpos),
                        (N.XR_BUF_MOTION, motion), (N.XR_BUF_SPATIAL, np.zeros(2 * n, N.RESERVOIR_DTYPE))]:
        ctx.write(which, data)
    p = N.default_params(temporal_m_cap=cap)
    p.passes = N.XR_PASS_TEMPORAL
    say("[cases] temporal pass on synthetic reservoirs, M cap 20")
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    sp = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    temporal(W, H, frame, cap, s, motion, gpos, t, sp, np.zeros(n, np.uint32))
    g = ctx.read(N.XR_BUF_TEMPORAL)
    eq = bytes_equal_rows(g, t)
    assert eq.all(), describe_mismatch("temporal reservoirs", g, t, eq)
    assert (g[n:]["m"] == cap).sum() > n // 4


def _spatial_on_device(scene, s, t, W, H, visibility):
    n = W * H
    ctx = N.Context(scene, W, H)
    view = look_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5))
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=2, reuse=1)
    ctx.write(N.XR_BUF_SAMPLES, s)
    ctx.write(N.XR_BUF_TEMPORAL, t)
    ctx.write(N.XR_BUF_SPATIAL, np.zeros(2 * n, N.RESERVOIR_DTYPE))
    p = N.default_params(spatial_samples=6, spatial_radius=30.0, spatial_m_cap=300, visibility=visibility)
    p.passes = N.XR_PASS_SPATIAL
    ctx.render(ubo, N.XR_MODE_RESTIR, p)
    ctx.sync()
    sp = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    oracle_scene(scene).spatial(W, H, 2, 1, 6, 30.0, 300, visibility, s, t, sp, np.zeros(n, np.uint32))
    g = ctx.read(N.XR_BUF_SPATIAL)
    eq = bytes_equal_rows(g, sp)
    assert eq.all(), describe_mismatch("spatial reservoirs", g, sp, eq)
    return g[n:]


@pytest.mark.parametrize("case", sorted(JACOBIAN_CASES))
def test_spatial_jacobian_on_device(case):
    xc, xo, xs, ns = JACOBIAN_CASES[case]
    W = H = 16
    c = 8 * W + 8
    s, t = spatial_layout(W, H, c, xc, xo, xc, xs, ns)
    say(f"[cases] spatial Jacobian, {case}")
    r = _spatial_on_device(cornell_scene(), s, t, W, H, N.XR_VIS_REFERENCE)[c]
    J = jacobian(xo, xc, xs, ns)
    cands = np.array([3.0 + a * 3.0 * J + (6 - a) * 3.0 for a in range(7)])
    k = int(np.argmin(np.abs(cands - r["w_sum"])))
    assert r["m"] == 14 and abs(cands[k] - r["w_sum"]) <= 2e-6 * cands[k] and k >= 4


@pytest.mark.parametrize("blocked", ["reservoir_origin", "pixel_origin"])
def test_spatial_visibility_origin_on_device(blocked):
    A, B = (-0.1, 0.0, 0.0), (0.1, 0.0, 0.0)
    res_xv, pix_xv = (A, B) if blocked == "reservoir_origin" else (B, A)
    W = H = 16
    s, t = spatial_layout(W, H, 8 * W + 8, pix_xv, pix_xv, res_xv, (0.0, 1.0, 0.0), (0.0, -1.0, 0.0))
    say(f"[cases] spatial visibility origin, {blocked} blocked")
    r = _spatial_on_device(occluder_scene(), s, t, W, H, N.XR_VIS_TRACE)
    assert np.all(r["m"] == (2 if blocked == "reservoir_origin" else 14))
