"""Closed-form pins of the oracle's pass arithmetic (no GPU).

Each test builds an input whose result the mathematics fixes, so that a
plausible restatement mistake fails it:
  * emitters shade with the untextured baseColor * 4 (InitialSample.rgen:117-120,
    rgen.hlsl:96-99) while the Lambertian albedo comes from the texture
    (BSDF.hlsl:22-23);
  * motion vectors (InitialSample.rgen:199-203, WorldToScreenUV Sampling.hlsl:39-56):
    a static camera leaves only the sub-pixel jitter, (0.5 - xi) / size; a
    camera translated by d perpendicular to its view, at depth 1, adds
    -0.5 d_x / (tan(fov/2) aspect) and +0.5 d_y / tan(fov/2) (v grows downwards);
  * temporal M cap (Temporal.rgen:100-129): w_sum = (p_new/p_q + p_prev W M) * cap / M
    and W = w_sum / (M p_hat(selected));
  * spatial reconnection Jacobian (Spatial.rgen:113-133) inside and at both
    clamps, and the visibility ray's origin, r_s.s.x_view (Spatial.rgen:135).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from tests.helpers import cornell_scene, look_view, oracle_scene
from tests.pass_cases import (JACOBIAN_CASES, floor_light_textured, jacobian, occluder_scene, spatial_layout,
                              translation_view, wall_scene)
from xvk_ray_amd import native as N


# ------------------------------------------------------------------ emitters ----
def test_emitter_ignores_its_texture_lambertian_uses_it():
    scene = floor_light_textured()
    orc = oracle_scene(scene)
    W, H = 16, 12
    view = look_view((0, 0.5, 0), (0.0, 0.0, -1.0))
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    s = orc.initial_sample(N.ubo_to_bytes(ubo), W, H)["samples"]
    floor = s["material_id"] == 0
    assert floor.sum() > W * H // 2
    # the light's texture is black: a textured emitter would give L_out = 0
    assert np.all(s["l_out"][floor] == np.float32(8.0))
    # the floor's texture is white: f = 1/pi, not base colour 0.5 / pi
    assert np.allclose(s["f"][floor], 1.0 / np.pi, rtol=1e-6)
    # plain path tracing at depth 2: floor -> light, radiance (1/pi * cos / (cos/pi)) * 8 = 8
    ubo = N.make_ubo(view, view, 60.0, W, H, spp=1, total=1, bounces=10, frame=1)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    orc.plain(N.ubo_to_bytes(ubo), W, H, 2, accum, out)
    assert np.allclose(accum[:, :3], 8.0, rtol=1e-5)


# -------------------------------------------------------------------- motion ----
def _jitter(oracle_lib, W, H, frame):
    """The first two RandomFloat draws of every pixel (the camera jitter)."""
    xi = np.zeros((W * H, 2), np.float64)
    words = np.zeros(2, np.uint32)
    fl = np.zeros(2, np.float32)
    for y in range(H):
        for x in range(W):
            oracle_lib.orc_random_stream(oracle_lib.orc_init_seed(x, y, W, frame), 2, words.ctypes.data,
                                         fl.ctypes.data)
            xi[y * W + x] = fl
    return xi


def _motion_f64(motion_u32):
    h = np.stack([motion_u32 & 0xFFFF, motion_u32 >> 16], 1).astype(np.uint16)
    return h.view(np.float16).astype(np.float64)


@pytest.mark.parametrize("delta", [(0.0, 0.0), (0.03, -0.02), (-0.05, 0.04)])
def test_motion_closed_form(oracle_lib, delta):
    W, H, fov, frame = 32, 24, 60.0, 3
    eye_prev = np.array([0.0, 0.0, 0.0])
    eye = eye_prev + [delta[0], delta[1], 0.0]
    ubo = N.make_ubo(translation_view(eye), translation_view(eye_prev), fov, W, H, bounces=1, frame=frame)
    out = oracle_scene(wall_scene()).initial_sample(N.ubo_to_bytes(ubo), W, H)
    assert np.all(out["gpos"][:, 3] == 1.0)  # every camera ray meets the wall
    mv = _motion_f64(out["motion"])
    xi = _jitter(oracle_lib, W, H, frame)
    tan_h = math.tan(math.radians(fov) / 2)
    want_x = (0.5 - xi[:, 0]) / W - 0.5 * delta[0] / (tan_h * W / H)
    want_y = (0.5 - xi[:, 1]) / H + 0.5 * delta[1] / tan_h
    # RG16F storage: half an ulp of the stored value plus the FP32 projection's error
    tol_x = np.abs(want_x) * 2.0 ** -11 + 2e-6
    tol_y = np.abs(want_y) * 2.0 ** -11 + 2e-6
    assert np.all(np.abs(mv[:, 0] - want_x) <= tol_x), np.abs(mv[:, 0] - want_x).max()
    assert np.all(np.abs(mv[:, 1] - want_y) <= tol_y), np.abs(mv[:, 1] - want_y).max()


# ------------------------------------------------------------------ temporal ----
def _target(s, x_curr, n_curr):
    """EvaluateTargetFunction (Temporal.rgen:33-40) in float64."""
    wi = s["x_sample"].astype(np.float64) - x_curr
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    n = n_curr / np.linalg.norm(n_curr, axis=1, keepdims=True)
    cos = np.sum(n * wi, axis=1)
    return np.where(cos > 0, np.linalg.norm(s["l_out"].astype(np.float64), axis=1), 0.0)


def test_temporal_cap_rescale_and_W():
    from oracle.binding import temporal
    from tests.test_cpu_oracle_passes import _synthetic_frame
    W, H, cap, frame = 40, 30, 20, 3  # frame 3: current slot n.., previous slot 0..n-1
    n = W * H
    s, t, gpos, motion, bad = _synthetic_frame(W, H, 2, invalid_frac=0.0)
    motion[:] = 0  # the previous pixel is the pixel itself
    rng = np.random.default_rng(4)
    t["s"][:n] = s[rng.permutation(n)]  # previous reservoirs hold other pixels' samples
    t["m"][:n] = rng.integers(0, 40, n)  # above and below the cap
    gpos[:, 3] = 1.0
    t0 = t.copy()
    out = np.zeros(n, np.uint32)
    temporal(W, H, frame, cap, s, motion, gpos, t, np.zeros(2 * n, N.RESERVOIR_DTYPE), out)
    cur, prev = t[n:], t0[:n]
    x_curr, n_curr = s["x_view"].astype(np.float64), s["n_view"].astype(np.float64)
    p_new = _target(s, x_curr, n_curr)
    p_prev = _target(prev["s"], x_curr, n_curr)
    w_sum = p_new / s["p_q"] + p_prev * prev["w"].astype(np.float64) * prev["m"]
    m = 1 + prev["m"].astype(np.int64)
    capped = m > cap
    assert capped.sum() > n // 4 and (~capped).sum() > n // 4
    w_sum = np.where(capped, w_sum * cap / m, w_sum)
    m = np.minimum(m, cap)
    assert np.array_equal(cur["m"], m)
    assert np.allclose(cur["w_sum"], w_sum, rtol=2e-6, atol=1e-30)
    # W from the selected sample's target at the current pixel
    took_new = np.all(cur["s"]["x_sample"] == s["x_sample"], axis=1)
    took_prev = np.all(cur["s"]["x_sample"] == prev["s"]["x_sample"], axis=1)
    # with both weights zero nothing is taken: the initial sample (L_out = 0) stays
    none = (p_new == 0) & (p_prev * prev["w"] * prev["m"] == 0)
    assert np.all(took_new | took_prev | none) and took_new.any() and took_prev.any()
    assert np.all(cur["s"]["l_out"][none & ~took_new & ~took_prev] == 0)
    p_sel = np.where(took_new, p_new, np.where(took_prev, p_prev, 0.0))
    W_want = np.where(p_sel > 1e-6, w_sum / (m * np.where(p_sel > 0, p_sel, 1.0)), 0.0)
    assert np.allclose(cur["w"], W_want, rtol=4e-6, atol=1e-30)


# ------------------------------------------------------------------- spatial ----
@pytest.mark.parametrize("case", sorted(JACOBIAN_CASES))
def test_spatial_jacobian_inside_and_at_both_clamps(case):
    """Centre pixel c has visible point C, every other pixel P; all reservoirs
    hold one sample (x_s, n_s, |L| = 3, W = 0.5, m = 2).  Each of the K = 6
    neighbours is another pixel (Jacobian J(P -> C)) or, rarely, c itself
    (Jacobian 1), so w_sum = 3 + a * 3J + (6 - a) * 3 for some a <= 6 and M = 14."""
    xc, xo, xs, ns = JACOBIAN_CASES[case]
    W = H = 16
    c = 8 * W + 8
    s, t = spatial_layout(W, H, c, xc, xo, xc, xs, ns)
    orc = oracle_scene(cornell_scene())
    sp = np.zeros_like(t)
    out = np.zeros(W * H, np.uint32)
    orc.spatial(W, H, 2, 1, 6, 30.0, 300, N.XR_VIS_REFERENCE, s, t, sp, out)
    r = sp[W * H + c]
    J = jacobian(xo, xc, xs, ns)
    assert {"inside": 0.1 < J < 10 and abs(J - 1) > 0.2, "low_clamp": J == 0.1, "high_clamp": J == 10.0}[case]
    assert r["m"] == 2 + 6 * 2
    cands = np.array([3.0 + a * 3.0 * J + (6 - a) * 3.0 for a in range(7)])
    k = int(np.argmin(np.abs(cands - r["w_sum"])))
    assert abs(cands[k] - r["w_sum"]) <= 2e-6 * cands[k], (r["w_sum"], cands)
    assert k >= 4  # self-reuse is rare: most neighbours are other pixels
    assert math.isclose(r["w"], r["w_sum"] / (14 * 3.0), rel_tol=2e-6)


@pytest.mark.parametrize("blocked", ["reservoir_origin", "pixel_origin"])
def test_spatial_visibility_ray_leaves_from_the_selected_sample(blocked):
    """The segment from (-0.1, 0, 0) to x_s = (0, 1, 0) is blocked, the one from
    (0.1, 0, 0) is not.  The visibility ray leaves from r_s.s.x_view, the
    reservoir sample's visible point, not from the pixel's own: with the
    reservoirs' x_view blocked every neighbour is dropped (M = 2, w_sum = 3);
    with the pixels' x_view blocked every neighbour merges (M = 14, w_sum = 21)."""
    A, B = (-0.1, 0.0, 0.0), (0.1, 0.0, 0.0)
    res_xv, pix_xv = (A, B) if blocked == "reservoir_origin" else (B, A)
    W = H = 16
    c = 8 * W + 8
    s, t = spatial_layout(W, H, c, pix_xv, pix_xv, res_xv, (0.0, 1.0, 0.0), (0.0, -1.0, 0.0))
    orc = oracle_scene(occluder_scene())
    sp = np.zeros_like(t)
    out = np.zeros(W * H, np.uint32)
    rays = orc.spatial(W, H, 2, 1, 6, 30.0, 300, N.XR_VIS_TRACE, s, t, sp, out)
    assert rays == 6 * W * H
    r = sp[W * H:]
    if blocked == "reservoir_origin":
        assert np.all(r["m"] == 2) and np.allclose(r["w_sum"], 3.0, rtol=1e-7)
    else:
        assert np.all(r["m"] == 14) and np.allclose(r["w_sum"], 21.0, rtol=2e-6)
