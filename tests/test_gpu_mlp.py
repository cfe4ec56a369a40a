"""MLP radiance cache on the GPU vs the CPU oracle, through the C ABI (needs an MI355X).

Tolerances: the kernels and the oracle's BF16 mode round at the same points
(inputs, hidden weights, hidden activations, backpropagated pre-activation
gradients to BF16; FP32 output layer), but the MFMA sums in an order of its
own, so a value near a BF16 rounding boundary can land one BF16 ulp (2^-8
relative) apart and carry that through later layers.  Forward outputs are
compared to 1e-2 of the output scale, gradients per layer to 5e-2 in relative
L2 norm with cosine similarity above 0.998.  The Adam step is pinned by its
closed form; path termination and training-pixel choice are compared bit for
bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
from tests.helpers import CORNELL_EYE, CORNELL_TARGET, bytes_equal_rows, describe_mismatch, look_view, oracle_scene, say
from tests.test_cpu_oracle_mlp import he_params, random_inputs
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _ctx(W=32, H=32):
    return N.Context(N.Scene.cornell(1), W, H)


def _layers():
    for l in range(6):
        wo, _ = B.mlp_layer_offsets(l)
        end = B.mlp_layer_offsets(l + 1)[0] if l < 5 else B.MLP_PARAMS
        yield l, wo, end


def _positive_params(seed):
    w = he_params(seed)
    w[B.mlp_layer_offsets(5)[1]:] = 5.0  # predictions away from 0 (see the oracle tests)
    return w


def test_params_roundtrip_and_init():
    ctx = _ctx()
    p0 = ctx.mlp_params()
    assert p0.shape == (B.MLP_PARAMS,) and np.isfinite(p0).all()
    for l, wo, end in _layers():
        bo = B.mlp_layer_offsets(l)[1]
        fan_in = 96 if l == 0 else 64
        a = np.sqrt(6.0 / fan_in)
        assert np.abs(p0[wo:bo]).max() <= a and np.abs(p0[wo:bo]).max() > 0.9 * a
        assert np.all(p0[bo:end] == 0.0)
    w = he_params(3)
    assert np.array_equal(ctx.mlp_params(w), w)
    ctx.mlp_reset(1)
    assert np.array_equal(ctx.mlp_params(), p0)
    ctx.mlp_reset(2)
    assert not np.array_equal(ctx.mlp_params(), p0)


@pytest.mark.parametrize("rows", [1, 63, 64, 1000, 20000])
def test_forward_matches_oracle(rows):
    ctx = _ctx()
    w = he_params(5)
    ctx.mlp_params(w)
    x = random_inputs(rows, 6 + rows)
    say(f"[mlp-eval] rows={rows}")
    gpu = ctx.mlp_eval(x).astype(np.float64)
    ref = B.mlp_forward(w, x, mode=1)
    exact = B.mlp_forward(w, x, mode=0)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(gpu - ref).max() <= 1e-2 * scale, float(np.abs(gpu - ref).max())
    assert np.abs(gpu - exact).max() <= 5e-2 * scale
    # most rows agree much closer than the bound
    assert np.median(np.abs(gpu - ref)) <= 1e-3 * scale


def test_forward_encoded_features():
    """Inputs of the shape the cache sees: encoded position, unit vectors, albedo."""
    ctx = _ctx()
    w = he_params(8)
    ctx.mlp_params(w)
    rng = np.random.default_rng(9)
    rows = 4096
    x = np.zeros((rows, 96), np.float32)
    lo = np.zeros(3, np.float32)
    inv = np.ones(3, np.float32)
    for r in range(rows):
        d = rng.normal(size=3)
        nn = rng.normal(size=3)
        x[r] = B.mlp_encode(rng.uniform(0, 1, 3), d / np.linalg.norm(d), nn / np.linalg.norm(nn),
                            rng.uniform(0, 1, 3), lo, inv)
    gpu = ctx.mlp_eval(x).astype(np.float64)
    ref = B.mlp_forward(w, x, mode=1)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(gpu - ref).max() <= 1e-2 * scale


# 40,001 rows: 626 tiles of 64 for at most 512 workgroups, so some workgroups
# loop to a second tile, and the last tile holds one row.
@pytest.mark.parametrize("rows", [64, 1000, 8192, 40001])
def test_gradient_matches_oracle(rows):
    ctx = _ctx()
    w = _positive_params(12)
    ctx.mlp_params(w)
    x = random_inputs(rows, 13)
    y = np.random.default_rng(14).uniform(0.0, 8.0, (rows, 3)).astype(np.float32)
    say(f"[mlp-grad] rows={rows}")
    g = ctx.mlp_gradient(x, y).astype(np.float64)
    _, ref = B.mlp_loss_grad(w, x, y, mode=1)
    for l, wo, end in _layers():
        a, b = g[wo:end], ref[wo:end]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel < 5e-2 and cos > 0.998, (l, rel, cos)
    # a gradient call takes no step
    assert np.array_equal(ctx.mlp_params(), w)


def test_adam_first_steps_closed_form():
    """Step 1 of bias-corrected Adam moves each weight by -lr * g / (|g| + eps);
    step 2 follows the oracle's Adam on the kernels' gradients."""
    ctx = _ctx()
    w0 = _positive_params(21)
    ctx.mlp_reset(1)  # zero Adam state and step count ...
    ctx.mlp_params(w0)  # ... with these weights
    rows = 4096
    x = random_inputs(rows, 22)
    y = np.random.default_rng(23).uniform(0.0, 8.0, (rows, 3)).astype(np.float32)
    lr = 1e-3
    g1 = ctx.mlp_gradient(x, y).astype(np.float64)
    ctx.mlp_train_rows(x, y, lr)
    w1 = ctx.mlp_params().astype(np.float64)
    step = w1 - w0.astype(np.float64)
    big = np.abs(g1) > 1e-3 * np.abs(g1).max()
    assert big.sum() > 1000
    assert np.allclose(step[big], -lr * np.sign(g1[big]), rtol=2e-2, atol=1e-6), \
        float(np.abs(step[big] + lr * np.sign(g1[big])).max())
    assert np.abs(step).max() <= lr * 1.001  # FP32 rounding of the weight (bias 5: ulp 4.8e-7)
    # second step against the oracle's Adam
    x2 = random_inputs(rows, 24)
    g2 = ctx.mlp_gradient(x2, y).astype(np.float64)
    ctx.mlp_train_rows(x2, y, lr)
    w2 = ctx.mlp_params().astype(np.float64)
    wr, m, v = w0.astype(np.float64), np.zeros(B.MLP_PARAMS), np.zeros(B.MLP_PARAMS)
    B.adam_step(wr, m, v, g1.copy(), lr, 0.9, 0.99, 1e-8, 1)
    B.adam_step(wr, m, v, g2.copy(), lr, 0.9, 0.99, 1e-8, 2)
    sel = (np.abs(g1) > 1e-3 * np.abs(g1).max()) & (np.abs(g2) > 1e-3 * np.abs(g2).max())
    d_gpu, d_ref = (w2 - w0)[sel], (wr - w0)[sel]
    assert np.abs(d_gpu - d_ref).max() <= 0.05 * lr


def test_training_reduces_loss():
    """A smooth radiance-like target of the encoded inputs is learned: the
    oracle's loss of the GPU-trained weights falls well below the start."""
    ctx = _ctx()
    ctx.mlp_reset(3)
    rng = np.random.default_rng(30)
    rows = 8192
    lo, inv = np.zeros(3, np.float32), np.ones(3, np.float32)
    pos = rng.uniform(0, 1, (rows, 3)).astype(np.float32)
    x = np.zeros((rows, 96), np.float32)
    nrm = np.array([0, 1, 0], np.float32)
    for r in range(rows):
        x[r] = B.mlp_encode(pos[r], nrm, nrm, np.full(3, 0.5, np.float32), lo, inv)
    y = np.stack([1.0 + np.sin(3 * pos[:, 0]), 0.5 + pos[:, 1], 2.0 * pos[:, 2] ** 2], 1).astype(np.float32)
    w_start = ctx.mlp_params()
    loss0, _ = B.mlp_loss_grad(w_start, x, y, mode=1, grad=False)
    say("[mlp-train] 300 steps")
    for _ in range(300):
        ctx.mlp_train_rows(x, y, 1e-3)
    w_end = ctx.mlp_params()
    loss1, _ = B.mlp_loss_grad(w_end, x, y, mode=1, grad=False)
    say(f"[mlp-train] loss {loss0:.4f} -> {loss1:.4f}")
    assert np.isfinite(w_end).all()
    assert loss1 < 0.01 * loss0 and loss1 < 0.1


# ---------------------------------------------------------------------------
# FP8 E4M3 weights
# ---------------------------------------------------------------------------
def test_fp8_codes_and_scales_bit_exact():
    ctx = _ctx()
    w = he_params(50)
    w[B.mlp_layer_offsets(2)[0]:B.mlp_layer_offsets(2)[1]] *= 3000.0
    ctx.mlp_params(w)
    with pytest.raises(N.XrError):
        ctx.mlp_fp8_weights()  # BF16 format: no FP8 copy
    ctx.mlp_set_weight_format(N.XR_MLP_WEIGHTS_FP8)
    codes, scales = ctx.mlp_fp8_weights()
    ref_codes, ref_scales = B.mlp_quantize(w)
    assert np.array_equal(scales, ref_scales)
    assert np.array_equal(codes, ref_codes), int((codes != ref_codes).sum())
    # an Adam step re-quantises from the updated master weights
    x = random_inputs(512, 51)
    y = np.ones((512, 3), np.float32)
    ctx.mlp_train_rows(x, y, 1e-3)
    codes2, scales2 = ctx.mlp_fp8_weights()
    rc2, rs2 = B.mlp_quantize(ctx.mlp_params())
    assert np.array_equal(codes2, rc2) and np.array_equal(scales2, rs2)
    assert not np.array_equal(codes2, codes)


@pytest.mark.parametrize("rows", [64, 5000])
def test_fp8_forward_and_gradient_match_oracle(rows):
    ctx = _ctx()
    w = _positive_params(52)
    ctx.mlp_params(w)
    ctx.mlp_set_weight_format(N.XR_MLP_WEIGHTS_FP8)
    x = random_inputs(rows, 53)
    say(f"[mlp-fp8] rows={rows}")
    gpu = ctx.mlp_eval(x).astype(np.float64)
    ref = B.mlp_forward(w, x, mode=2)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(gpu - ref).max() <= 1e-2 * scale, float(np.abs(gpu - ref).max())
    bf = B.mlp_forward(w, x, mode=1)
    assert np.abs(gpu - bf).max() > np.abs(gpu - ref).max()  # the FP8 weights are in use
    y = np.random.default_rng(54).uniform(0.0, 8.0, (rows, 3)).astype(np.float32)
    g = ctx.mlp_gradient(x, y).astype(np.float64)
    _, gr = B.mlp_loss_grad(w, x, y, mode=2)
    for l, wo, end in _layers():
        a, b = g[wo:end], gr[wo:end]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel < 5e-2 and cos > 0.998, (l, rel, cos)


def test_fp8_training_reduces_loss():
    ctx = _ctx()
    ctx.mlp_reset(3)
    ctx.mlp_set_weight_format(N.XR_MLP_WEIGHTS_FP8)
    rng = np.random.default_rng(60)
    rows = 8192
    lo, inv = np.zeros(3, np.float32), np.ones(3, np.float32)
    pos = rng.uniform(0, 1, (rows, 3)).astype(np.float32)
    nrm = np.array([0, 1, 0], np.float32)
    x = np.stack([B.mlp_encode(p, nrm, nrm, np.full(3, 0.5, np.float32), lo, inv) for p in pos])
    y = np.stack([1.0 + np.sin(3 * pos[:, 0]), 0.5 + pos[:, 1], 2.0 * pos[:, 2] ** 2], 1).astype(np.float32)
    loss0, _ = B.mlp_loss_grad(ctx.mlp_params(), x, y, mode=2, grad=False)
    for _ in range(300):
        ctx.mlp_train_rows(x, y, 1e-3)
    loss1, _ = B.mlp_loss_grad(ctx.mlp_params(), x, y, mode=2, grad=False)
    say(f"[mlp-fp8-train] loss {loss0:.4f} -> {loss1:.4f}")
    assert loss1 < 0.01 * loss0 and loss1 < 0.2


def _cache_frame(ctx, W, H, frame, params):
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    ubo = N.make_ubo(view, view, 45.0, W, H, bounces=4, frame=frame)
    ctx.render(ubo, N.XR_MODE_RESTIR, params)
    ctx.sync()
    return ubo


def _features(samples, mats, lo, inv, idx):
    f32 = np.float32
    x = np.zeros((len(idx), 96), np.float32)
    for r, p in enumerate(idx):
        s = samples[p]
        d = (s["x_view"] - s["x_sample"]).astype(np.float32)
        dd = f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2])
        wv = (d * (f32(1.0) / np.sqrt(f32(dd)))).astype(np.float32)
        alb = mats["base_color"][s["pad1"]][:3]
        x[r] = B.mlp_encode(s["x_sample"], wv, s["n_sample"], alb, lo, inv)
    return x


@pytest.mark.parametrize("fp8", [False, True])
def test_cache_frame_matches_oracle(fp8):
    """One cache frame: termination flags and every traced field bit-exact vs
    the oracle; query and training pixels carry max(0, network(features))
    computed with the weights after the frame's training step (BF16 or FP8
    weights)."""
    scene = N.Scene.cornell(1)
    W, H = 64, 48
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    verts, _, mats, _ = scene.arrays()
    lo = verts["pos"].min(0).astype(np.float32)
    inv = (np.float32(1.0) / (verts["pos"].max(0) - lo).astype(np.float32)).astype(np.float32)
    stride = 8
    flags = N.XR_FLAG_MLP_CACHE | (N.XR_FLAG_MLP_FP8 if fp8 else 0)
    params = N.default_params(flags=flags, mlp_train_pixels=W * H // stride, mlp_spread=0.01)
    # train a few frames first so the network is not at its initial state
    for frame in range(1, 4):
        _cache_frame(ctx, W, H, frame, params)
    w_before = ctx.mlp_params()
    frame = 4
    say(f"[cache-frame] {W}x{H} frame {frame}")
    ctx.reset_counters()
    ubo = _cache_frame(ctx, W, H, frame, params)
    counts = ctx.counters_all()
    w_after = ctx.mlp_params()
    assert not np.array_equal(w_before, w_after)  # the frame trained the network
    gpu = ctx.read(N.XR_BUF_SAMPLES)
    T = W * H // stride
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, cache=(T, 0.01))["samples"]
    kind = ref["pad2"]
    assert np.array_equal(gpu["pad2"], kind) and np.array_equal(gpu["pad1"], ref["pad1"])
    query = kind == 1
    assert query.sum() > 0.1 * W * H and (kind == 2).sum() > 0
    assert counts[3] == query.sum() and counts[4] == (kind == 2).sum() == T  # more candidates than T here
    # everything but the cached radiance is bit-exact
    train = kind == 2
    g2, r2 = gpu.copy(), ref.copy()
    g2["l_out"][query | train] = 0
    r2["l_out"][train] = 0  # the traced suffix radiance, consumed by the training step
    eq = bytes_equal_rows(g2, r2)
    assert eq.all(), describe_mismatch("cache-frame samples", g2, r2, eq)
    # query and training pixels carry the network's radiance after the step
    idx = np.nonzero(query | train)[0]
    pred = np.maximum(B.mlp_forward(w_after, _features(ref, mats, lo, inv, idx), mode=2 if fp8 else 1), 0.0)
    got = gpu["l_out"][idx].astype(np.float64)
    scale = max(1.0, np.abs(pred).max())
    assert np.abs(got - pred).max() <= 1e-2 * scale, float(np.abs(got - pred).max())


def test_cache_frame_training_step_matches_oracle_gradient():
    """The frame-mode training input (features encoded per wave from the
    sample planes, targets = traced suffix radiance of the listed training
    pixels): the first Adam step after a reset moves every weight by
    -lr * sign(g), so its signs must be the signs of the oracle's gradient of
    the same loss on the oracle's features and traced targets of the training
    pixels.  A wrong column, axis, albedo or target mapping scrambles them."""
    scene = N.Scene.cornell(1)
    W, H = 64, 48
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    verts, _, mats, _ = scene.arrays()
    lo = verts["pos"].min(0).astype(np.float32)
    inv = (np.float32(1.0) / (verts["pos"].max(0) - lo).astype(np.float32)).astype(np.float32)
    stride, lr = 4, 1e-3
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE, mlp_train_pixels=W * H // stride, mlp_spread=0.01,
                              mlp_lr=lr)
    ctx.mlp_reset(5)
    w0 = ctx.mlp_params()
    w0[B.mlp_layer_offsets(5)[1]:] = 1.0  # output biases away from 0: no dead outputs on the first step
    ctx.mlp_params(w0)
    say(f"[cache-train] first cache frame {W}x{H}, stride {stride}")
    ubo = _cache_frame(ctx, W, H, 1, params)
    w1 = ctx.mlp_params().astype(np.float64)
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, cache=(W * H // stride, 0.01))["samples"]
    train = np.nonzero(ref["pad2"] == 2)[0]
    assert len(train) > 200
    x = _features(ref, mats, lo, inv, train)
    y = ref["l_out"][train].astype(np.float32)
    _, g = B.mlp_loss_grad(w0, x, y, mode=1)
    step = w1 - w0.astype(np.float64)
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    huge = np.abs(g) > 1e-1 * np.abs(g).max()
    assert big.sum() > 500
    agree = np.sign(step[big]) == -np.sign(g[big])
    assert agree.mean() >= 0.995, agree.mean()
    assert np.all(np.sign(step[huge]) == -np.sign(g[huge]))
    assert np.allclose(np.abs(step[big]), lr, rtol=2e-2)


def test_cache_gain_over_plain_restir_and_untrained():
    """What the trained cache gains over the method without it, at equal
    spatial neighbours (5): on the Cornell box at 256x192 with the reference's
    10 bounces, 128 frames, its radiance is closer to the converged image than
    plain ReSTIR GI's and than the untrained cache's (lr 0), and it traces
    fewer extension rays than plain ReSTIR GI."""
    scene = N.Scene.cornell(1)
    W, H = 256, 192
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    frames, bounces = 128, 10

    ctx = N.Context(scene, W, H)

    def run(flags, lr):
        ctx.resize(W, H)  # empty history (and converged sums)
        ctx.converge(N.make_ubo(view, view, 45.0, W, H, bounces=bounces, frame=1), 1024, 1 << 20)
        ctx.mlp_reset(1)
        p = N.default_params(flags=N.XR_FLAG_RADIANCE | flags, mlp_train_pixels=4096, mlp_lr=lr,
                             visibility=N.XR_VIS_TRACE, spatial_samples=5, spatial_radius=30.0)
        errs = []
        ctx.reset_counters()
        ctx.reset_pass_times()
        for f in range(1, frames + 1):
            ctx.render(N.make_ubo(view, view, 45.0, W, H, bounces=bounces, frame=f), N.XR_MODE_RESTIR, p)
            if f > frames - 8:
                ctx.sync()
                errs.append(ctx.rmse())
        ms, nfr = ctx.pass_time_totals()
        ext, _, _ = ctx.counters3()
        return float(np.mean(errs)), ext // nfr, sum(ms) / nfr

    say("[cache-gain] plain / untrained / trained")
    e_plain, r_plain, t_plain = run(0, 1e-3)
    e_untrained, r_untrained, _ = run(N.XR_FLAG_MLP_CACHE, 0.0)
    e_trained, r_trained, t_trained = run(N.XR_FLAG_MLP_CACHE, 1e-3)
    say(f"[cache-gain] rmse plain {e_plain:.4f} untrained {e_untrained:.4f} trained {e_trained:.4f}; "
        f"extend rays/frame {r_plain} {r_untrained} {r_trained}; ms/frame plain {t_plain:.3f} cache {t_trained:.3f}")
    assert e_trained < 0.98 * e_plain  # measured 0.0711 against 0.0750
    assert e_trained < 0.7 * e_untrained
    assert r_trained < 0.8 * r_plain and r_untrained < r_plain


@pytest.mark.parametrize("size", [(37, 11), (1, 1)])
def test_cache_frame_odd_size_fewer_candidates_than_target(size):
    """A cache frame on an image smaller than a tile and the training choice's
    4,096-pixel block, with a training target above the frame's candidate
    count: every candidate trains, termination flags and traced fields are
    bit-exact against the oracle, training pixels carry the network's
    radiance after the step."""
    W, H = size
    scene = N.Scene.cornell(1)
    ctx = N.Context(scene, W, H)
    orc = oracle_scene(scene)
    verts, _, mats, _ = scene.arrays()
    lo = verts["pos"].min(0).astype(np.float32)
    inv = (np.float32(1.0) / (verts["pos"].max(0) - lo).astype(np.float32)).astype(np.float32)
    T = 65536
    params = N.default_params(flags=N.XR_FLAG_MLP_CACHE, mlp_train_pixels=T, mlp_spread=0.01)
    for frame in (1, 2):
        _cache_frame(ctx, W, H, frame, params)
    ctx.reset_counters()
    ubo = _cache_frame(ctx, W, H, 3, params)
    counts = ctx.counters_all()
    w_after = ctx.mlp_params()
    gpu = ctx.read(N.XR_BUF_SAMPLES)
    ref = orc.initial_sample(N.ubo_to_bytes(ubo), W, H, cache=(T, 0.01))["samples"]
    kind = ref["pad2"]
    assert np.array_equal(gpu["pad2"], kind) and np.array_equal(gpu["pad1"], ref["pad1"])
    train = kind == 2
    assert not (kind == 1).any()  # T exceeds the candidates: none stops into the cache
    assert counts[3] == 0 and counts[4] == train.sum()
    g2, r2 = gpu.copy(), ref.copy()
    g2["l_out"][train] = 0
    r2["l_out"][train] = 0
    eq = bytes_equal_rows(g2, r2)
    assert eq.all(), describe_mismatch("cache-frame samples", g2, r2, eq)
    if train.any():
        idx = np.nonzero(train)[0]
        pred = np.maximum(B.mlp_forward(w_after, _features(ref, mats, lo, inv, idx), mode=1), 0.0)
        got = gpu["l_out"][idx].astype(np.float64)
        assert np.abs(got - pred).max() <= 1e-2 * max(1.0, np.abs(pred).max())
