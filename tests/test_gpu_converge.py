"""FP64 converged image and RMSE reduction vs the oracle (needs an MI355X).

Each converged sample is the initial sample's one-sample estimate
L_out * (1/p_q) * f * max(0, n_v . normalize(x_s - x_v)), formed in FP32 in
that order and summed in FP64 in sample order; the oracle's initial-sample
pass gives the same samples bit for bit, so the sums must agree exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import CORNELL_EYE, CORNELL_TARGET, look_view, oracle_scene, say
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _estimate(s: np.ndarray) -> np.ndarray:
    """FP32 per-sample estimate from reference-layout samples (numpy float32, no FMA)."""
    f32 = np.float32
    valid = (s["material_id"] != 0xFFFFFFFF) & (s["p_q"] > f32(1e-6))
    d = s["x_sample"] - s["x_view"]
    dd = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    dd = np.where(valid, dd, f32(1.0))  # invalid pixels are zeroed below
    inv = f32(1.0) / np.sqrt(dd)
    wi = d * inv[:, None]
    nv = s["n_view"]
    c = np.maximum(f32(0.0), nv[:, 0] * wi[:, 0] + nv[:, 1] * wi[:, 1] + nv[:, 2] * wi[:, 2])
    W = f32(1.0) / np.where(valid, s["p_q"], f32(1.0))
    out = ((s["l_out"] * W[:, None]) * s["f"]) * c[:, None]
    out[~valid] = 0
    return out.astype(np.float32)


def test_converged_sums_match_oracle():
    scene = N.Scene.cornell(1)
    W, H = 48, 32
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    spp, base = 6, 1000
    say(f"[converge] {W}x{H} spp={spp}")
    ctx.converge(ubo, spp, base)
    ctx.sync()
    gpu = ctx.converged_mean() * spp
    ref = np.zeros((W * H, 3), np.float64)
    for s in range(spp):
        u = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=base + s)
        samples = orc.initial_sample(N.ubo_to_bytes(u), W, H)["samples"]
        ref += _estimate(samples).astype(np.float64)
    assert np.array_equal(gpu, ref), float(np.abs(gpu - ref).max())
    assert ctx.converged_spp() == spp


def test_rmse_reduction_and_convergence():
    scene = N.Scene.cornell(1)
    W, H = 64, 48
    ctx = N.Context(scene, W, H)
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    ubo0 = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    ctx.converge(ubo0, 256, 100000)
    params = N.default_params(flags=N.XR_FLAG_RADIANCE)
    errs = []
    for frame in range(1, 9):
        ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=frame)
        ctx.render(ubo, N.XR_MODE_RESTIR, params)
        ctx.sync()
        rad = ctx.read(N.XR_BUF_RADIANCE)[:, :3].astype(np.float64)
        mean = ctx.converged_mean()
        host = float(np.sqrt(np.mean((rad - mean) ** 2)))
        dev = ctx.rmse()
        assert abs(dev - host) <= 1e-9 * max(1.0, host), (dev, host)
        errs.append(dev)
    say(f"[rmse] {['%.4f' % e for e in errs]}")
    # temporal reuse accumulates history: later frames are no worse than the first
    assert min(errs[4:]) < errs[0]
