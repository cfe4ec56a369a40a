"""xr_mlp_gradient_noalloc against xr_mlp_gradient (needs an MI355X).

The no-allocation entry point takes device rows and writes a device
gradient, keeping its partial sums in the library's static scratch array.
It runs the same kernels as xr_mlp_gradient, whose partial copies are summed
in the same order in FP64; only the order of the FP32 atomic additions inside
a copy may differ between two launches, so the gradients agree to 1e-5 of
each layer's largest entry.  Two calls in a row, the second with fewer rows,
check that the shared array is cleared between calls.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import binding as B
from tests.helpers import say
from tests.test_cpu_oracle_mlp import he_params, random_inputs
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def _noalloc(ctx, x, y):
    import torch
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    yd = torch.from_numpy(np.ascontiguousarray(y, np.float32)).to(dev)
    gd = torch.full((B.MLP_PARAMS,), float("nan"), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    rc = N.lib().xr_mlp_gradient_noalloc(ctx._h, C.c_void_p(xd.data_ptr()), C.c_void_p(yd.data_ptr()), len(x),
                                         C.c_void_p(gd.data_ptr()))
    assert rc == 0, N.lib().xr_last_error(ctx._h)
    ctx.sync()
    return gd.cpu().numpy()


def _close(g, ref):
    for l in range(6):
        wo, _ = B.mlp_layer_offsets(l)
        end = B.mlp_layer_offsets(l + 1)[0] if l < 5 else B.MLP_PARAMS
        a, b = g[wo:end].astype(np.float64), ref[wo:end].astype(np.float64)
        scale = max(np.abs(b).max(), 1e-30)
        assert np.abs(a - b).max() <= 1e-5 * scale, (l, np.abs(a - b).max(), scale)


def test_noalloc_gradient_matches_allocating_entry_point():
    ctx = N.Context(N.Scene.cornell(1), 16, 16)
    w = he_params(21)
    w[B.mlp_layer_offsets(5)[1]:] = 5.0
    ctx.mlp_params(w)
    for rows, seed in [(5000, 31), (700, 32)]:  # second call smaller: stale partials would show
        x = random_inputs(rows, seed)
        y = np.random.default_rng(seed + 100).uniform(0.0, 8.0, (rows, 3)).astype(np.float32)
        say(f"[noalloc] gradient of {rows} rows")
        ref = ctx.mlp_gradient(x, y)
        g = _noalloc(ctx, x, y)
        assert np.isfinite(g).all() and np.abs(g).max() > 0
        _close(g, ref)


def test_noalloc_gradient_refuses_without_network_or_beyond_maximum():
    ctx = N.Context(N.Scene.cornell(1), 16, 16)
    L = N.lib()
    p = C.c_void_p(16)  # never dereferenced: both calls fail their host-side checks
    assert L.xr_mlp_gradient_noalloc(ctx._h, p, p, 64, p) == N.XR_ERR_STATE
    ctx.mlp_reset(1)
    assert L.xr_mlp_gradient_noalloc(ctx._h, p, p, 4194305, p) == N.XR_ERR_INVALID_ARGUMENT
    assert b"maximum" in L.xr_last_error(ctx._h)
    assert L.xr_mlp_gradient_noalloc(ctx._h, None, p, 64, p) == N.XR_ERR_INVALID_ARGUMENT
