"""CPU tests pinning the oracle of the MLP radiance cache (no GPU).

The cache has no reference counterpart, so the oracle is pinned by what the
mathematics fixes:
  * BF16 rounding equals torch's FP32 -> bfloat16 conversion (round to
    nearest even), ties included;
  * the input encoding equals sin / cos of the same FP32 arguments in FP64;
  * the exact (FP64) network equals a matrix-form numpy evaluation, and its
    analytic gradient equals central finite differences of the loss with the
    relative-L2 denominator held fixed;
  * the BF16-emulating forward / backward stay within BF16 distance of the
    exact ones;
  * Adam: with a constant gradient every bias-corrected step is
    -lr * g / (|g| + eps), a zero gradient leaves the weights alone;
  * cache path termination: spread threshold 0 turns every non-training
    secondary diffuse vertex into a query (L = 0, path cut), an infinite
    threshold turns none, and training pixels equal the plain path.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import binding as B
from tests.helpers import CORNELL_EYE, CORNELL_TARGET, cornell_scene, look_view, oracle_scene
from xvk_ray_amd import native as N


def he_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    w = np.zeros(B.MLP_PARAMS, np.float32)
    for l in range(6):
        fan_in = 96 if l == 0 else 64
        fan_out = 3 if l == 5 else 64
        wo, bo = B.mlp_layer_offsets(l)
        a = np.sqrt(6.0 / fan_in)
        w[wo:wo + fan_in * fan_out] = rng.uniform(-a, a, fan_in * fan_out)
        w[bo:bo + fan_out] = rng.uniform(-0.05, 0.05, fan_out)
    return w


def random_inputs(rows: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = np.zeros((rows, 96), np.float32)
    x[:, :81] = rng.uniform(-1, 1, (rows, 81))
    return x


def numpy_forward(w: np.ndarray, x: np.ndarray) -> np.ndarray:
    h = x.astype(np.float64)
    for l in range(6):
        fi = 96 if l == 0 else 64
        fo = 3 if l == 5 else 64
        wo, bo = B.mlp_layer_offsets(l)
        W = w[wo:wo + fi * fo].astype(np.float64).reshape(fi, fo)
        h = h @ W + w[bo:bo + fo]
        if l < 5:
            h = np.maximum(h, 0.0)
    return h


def test_layout_offsets():
    assert B.Oracle.lib().orc_mlp_param_count() == 23043 == B.MLP_PARAMS
    assert B.mlp_layer_offsets(0) == (0, 96 * 64)
    assert B.mlp_layer_offsets(1) == (96 * 64 + 64, 96 * 64 + 64 + 4096)
    assert B.mlp_layer_offsets(5) == (23043 - 195, 23043 - 3)


def test_bf16_matches_torch_rne():
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.normal(size=20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 20000),
        # exact ties between two BF16 neighbours: odd and even lower neighbours
        (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        -(np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, 1.0, 3.0e38, -3.0e38, 1e-40], np.float32),
    ]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = B.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_e4m3_matches_torch_rne():
    rng = np.random.default_rng(40)
    finite = torch.arange(0, 127, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float32).numpy()
    mids = ((finite[1:] + finite[:-1]) / 2).astype(np.float32)  # exact ties
    x = np.concatenate([rng.uniform(-448, 448, 20000), rng.normal(size=20000) * 10.0 ** rng.integers(-4, 2, 20000),
                        finite, -finite, mids, -mids, np.nextafter(mids, 0), np.nextafter(mids, 1000)])
    x = np.clip(x, -448, 448).astype(np.float32)
    codes, vals = B.e4m3_round(x)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn)
    assert np.array_equal(codes, ref.view(torch.uint8).numpy())
    assert np.array_equal(vals, ref.to(torch.float32).numpy())
    # saturation (torch makes NaN beyond 448; the kernels clamp)
    c, v = B.e4m3_round(np.array([449.0, 1e6, -1e30], np.float32))
    assert list(v) == [448.0, 448.0, -448.0] and list(c) == [0x7E, 0x7E, 0xFE]


def test_fp8_layer_quantisation():
    w = he_params(41)
    w[B.mlp_layer_offsets(2)[0]:B.mlp_layer_offsets(2)[1]] *= 3000.0  # a layer that needs a scale > 1
    codes, scales = B.mlp_quantize(w)
    off = 0
    for l in range(5):
        wo, bo = B.mlp_layer_offsets(l)
        fi = 96 if l == 0 else 64
        W = w[wo:bo].reshape(fi, 64)
        s = float(scales[l])
        assert s == 2.0 ** round(np.log2(s))  # power of two
        assert np.abs(W).max() / s <= 448.0 < np.abs(W).max() / (s / 2)
        c = codes[off:off + 64 * fi].reshape(64, fi)  # W^T
        ref_codes, ref_vals = B.e4m3_round((W / np.float32(s)).astype(np.float32).T)
        assert np.array_equal(c.reshape(-1), ref_codes)
        # quantisation error: at most half an E4M3 ulp (2^-4 relative) above the subnormal range
        deq = ref_vals.reshape(64, fi).T * s
        big = np.abs(W) > s * 2.0 ** -6
        assert np.all(np.abs(deq - W)[big] <= np.abs(W)[big] * 2.0 ** -4 + 1e-12)
        off += 64 * fi
    assert off == 22528


def test_fp8_forward_close_to_bf16():
    w = he_params(42)
    x = random_inputs(500, 43)
    b = B.mlp_forward(w, x, mode=1)
    f = B.mlp_forward(w, x, mode=2)
    err = np.abs(f - b).max() / np.abs(b).max()
    assert 1e-3 < err < 0.2


def test_encoding_layout():
    p = np.array([0.3, -1.7, 2.2], np.float32)
    lo = np.array([-2.0, -3.0, 0.5], np.float32)
    ext = np.array([4.0, 3.5, 2.0], np.float32)
    inv = (np.float32(1.0) / ext).astype(np.float32)
    w = np.array([0.0, 0.6, 0.8], np.float32)
    n = np.array([1.0, 0.0, 0.0], np.float32)
    a = np.array([0.1, 0.2, 0.3], np.float32)
    f = B.mlp_encode(p, w, n, a, lo, inv)
    u = ((p - lo).astype(np.float32) * inv).astype(np.float32)
    for ax in range(3):
        scale = np.float32(3.14159265359)
        for k in range(12):
            arg = np.float32(scale * u[ax])
            assert abs(f[ax * 24 + 2 * k] - np.sin(np.float64(arg))) < 2e-6
            assert abs(f[ax * 24 + 2 * k + 1] - np.cos(np.float64(arg))) < 2e-6
            scale = np.float32(scale * np.float32(2.0))
    assert np.array_equal(f[72:75], w) and np.array_equal(f[75:78], n) and np.array_equal(f[78:81], a)
    assert np.all(f[81:] == 0.0)


def test_exact_forward_matches_matrix_form():
    w = he_params(1)
    x = random_inputs(300, 2)
    got = B.mlp_forward(w, x, mode=0)
    assert np.allclose(got, numpy_forward(w, x), rtol=1e-12, atol=1e-12)


def test_bf16_forward_close_to_exact():
    w = he_params(3)
    x = random_inputs(500, 4)
    exact = B.mlp_forward(w, x, mode=0)
    emu = B.mlp_forward(w, x, mode=1)
    err = np.abs(emu - exact).max() / np.abs(exact).max()
    assert 0 < err < 3e-2


def test_gradient_matches_finite_differences():
    w = he_params(7)
    rows = 64
    x = random_inputs(rows, 8)
    y = np.random.default_rng(9).uniform(0.0, 2.0, (rows, 3)).astype(np.float32)
    _, g = B.mlp_loss_grad(w, x, y, mode=0)
    p0 = B.mlp_forward(w, x, mode=0)
    den = p0 * p0 + 0.01  # the denominator is held constant (stop-gradient)

    def loss(wv):
        p = B.mlp_forward(wv, x, mode=0)
        return np.sum((p - y) ** 2 / den) / (3 * rows)

    rng = np.random.default_rng(10)
    idx = []
    for l in range(6):
        wo, bo = B.mlp_layer_offsets(l)
        idx += list(rng.integers(wo, bo, 12)) + list(rng.integers(bo, bo + (3 if l == 5 else 64), 3))
    checked = 0
    for i in idx:
        h = max(abs(float(w[i])) * 1e-3, 1e-4)
        wp, wm = w.copy(), w.copy()
        wp[i] = np.float32(w[i] + h)
        wm[i] = np.float32(w[i] - h)
        step = float(wp[i]) - float(wm[i])
        fd = (loss(wp) - loss(wm)) / step
        assert abs(fd - g[i]) <= 1e-4 * max(abs(fd), abs(g[i])) + 1e-9, (i, fd, g[i])
        checked += 1
    assert checked == len(idx)
    # the gradient is not trivially zero in any layer
    for l in range(6):
        wo, bo = B.mlp_layer_offsets(l)
        assert np.abs(g[wo:bo]).max() > 0


def test_bf16_gradient_close_to_exact():
    # Output bias 5 keeps predictions away from 0, where the relative-L2
    # weight 1 / (p^2 + 0.01) would magnify BF16 differences a hundredfold.
    w = he_params(11)
    w[B.mlp_layer_offsets(5)[1]:] = 5.0
    x = random_inputs(1024, 12)
    y = np.random.default_rng(13).uniform(0.0, 8.0, (1024, 3)).astype(np.float32)
    l0, g0 = B.mlp_loss_grad(w, x, y, mode=0)
    l1, g1 = B.mlp_loss_grad(w, x, y, mode=1)
    assert abs(l1 - l0) < 1e-2 * l0
    for l in range(6):
        wo, _ = B.mlp_layer_offsets(l)
        end = B.mlp_layer_offsets(l + 1)[0] if l < 5 else B.MLP_PARAMS
        a, b = g1[wo:end], g0[wo:end]
        # per-row terms partly cancel in the sum over rows, so the BF16
        # rounding of the backpropagated terms shows up enlarged in layer 0
        assert np.linalg.norm(a - b) < 0.1 * np.linalg.norm(b), l
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.99, l


def test_loss_closed_forms():
    w = he_params(14)
    x = random_inputs(32, 15)
    p = B.mlp_forward(w, x, mode=0).astype(np.float32)
    loss, g = B.mlp_loss_grad(w, x, p, mode=0)  # targets = predictions (to FP32)
    assert loss < 1e-10 and np.abs(g).max() < 1e-5
    y = np.zeros_like(p)
    loss, _ = B.mlp_loss_grad(w, x, y, mode=0)
    pd = B.mlp_forward(w, x, mode=0)
    assert loss == pytest.approx(np.mean(pd * pd / (pd * pd + 0.01)), rel=1e-12)


def test_adam_closed_forms():
    rng = np.random.default_rng(16)
    n = 1000
    g = rng.normal(size=n) * 10.0 ** rng.integers(-6, 2, n)
    lr, b1, b2, eps = 1e-3, 0.9, 0.99, 1e-8
    w = rng.normal(size=n)
    w0 = w.copy()
    m = np.zeros(n)
    v = np.zeros(n)
    for t in range(1, 6):
        before = w.copy()
        B.adam_step(w, m, v, g.copy(), lr, b1, b2, eps, t)
        # constant gradient: bias-corrected moments are exactly g and g^2
        assert np.allclose(before - w, lr * g / (np.abs(g) + eps), rtol=1e-9, atol=1e-15)
    assert np.allclose(w0 - w, 5 * lr * g / (np.abs(g) + eps), rtol=1e-9)
    # zero gradient after zero state: nothing moves
    w2, m2, v2 = w0.copy(), np.zeros(n), np.zeros(n)
    B.adam_step(w2, m2, v2, np.zeros(n), lr, b1, b2, eps, 1)
    assert np.array_equal(w2, w0)


def _cornell_ubo(W, H, frame=3):
    view = look_view(CORNELL_EYE, CORNELL_TARGET)
    return N.make_ubo(view, view, 45.0, W, H, bounces=4, frame=frame)


def test_cache_termination_invariants():
    W, H = 48, 40
    orc = oracle_scene(cornell_scene())
    ub = N.ubo_to_bytes(_cornell_ubo(W, H))
    plain = orc.initial_sample(ub, W, H)
    T = 97
    always = orc.initial_sample(ub, W, H, cache=(T, 0.0))
    never = orc.initial_sample(ub, W, H, cache=(T, 1e30))
    s_p, s_a, s_n = plain["samples"], always["samples"], never["samples"]
    train = always["train"].astype(bool)
    assert np.array_equal(train, never["train"].astype(bool))  # the choice does not depend on the threshold
    kind_a, kind_n = s_a["pad2"], s_n["pad2"]
    # training pixels: flagged, path identical to the plain one
    fields = ["x_view", "n_view", "x_sample", "n_sample", "l_out", "f", "p_q", "material_id", "rand_seed"]
    for k in fields:
        assert np.array_equal(s_a[k][train], s_p[k][train]), k
        assert np.array_equal(s_n[k], s_p[k]), k  # infinite threshold: no path is cut
    assert np.array_equal(kind_a == 2, train) and np.array_equal(kind_n == 2, train)
    assert set(np.unique(kind_n)) <= {0, 2}
    # threshold 0: every other pixel whose path reached a diffuse x_s is a query with L = 0
    query = kind_a == 1
    assert query.sum() > 0.3 * W * H
    assert np.all(~train[query])
    assert np.all(s_a["l_out"][query] == 0.0)
    assert np.array_equal(s_a["x_sample"][query], s_p["x_sample"][query])
    # pad1 names a non-emissive material at x_s
    mats = cornell_scene().arrays()[2]
    assert np.all(mats["material_model"][s_a["pad1"][query]] == N.MATERIAL_LAMBERTIAN)
    assert always["rays"] < plain["rays"] == never["rays"]
    # no emitter or miss at x_s becomes a query: those pixels equal the plain path
    rest = (kind_a == 0) & ~train
    for k in fields:
        assert np.array_equal(s_a[k][rest], s_p[k][rest]), k


def _gaps(chosen_ranks, C):
    r = np.sort(chosen_ranks)
    return np.diff(np.concatenate([r, [r[0] + C]]))


@pytest.mark.parametrize("T", [1, 64, 97, 500, 100000])
def test_cache_training_choice_exact_count_even_spacing(T):
    """Of C candidates (diffuse x_s), exactly min(T, C) train, spread evenly
    over the candidates in pixel order (cyclic gaps floor(C/T) or ceil(C/T)),
    and the choice moves with the frame."""
    W, H = 48, 40
    orc = oracle_scene(cornell_scene())
    sets = []
    for frame in (3, 4):
        out = orc.initial_sample(N.ubo_to_bytes(_cornell_ubo(W, H, frame)), W, H, cache=(T, 0.0))
        kind, train = out["samples"]["pad2"], out["train"].astype(bool)
        cand = (kind == 1) | (kind == 2)
        C = int(cand.sum())
        assert C > 0.3 * W * H and np.all(cand[train])
        assert train.sum() == min(T, C)
        if T < C:
            ranks = np.cumsum(cand)[train] - 1
            g = _gaps(ranks, C)
            assert set(np.unique(g)) <= {C // T, -(-C // T)}, np.unique(g)
        sets.append(train)
    if T < 500:
        assert not np.array_equal(sets[0], sets[1])
