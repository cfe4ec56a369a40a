"""Resources of the frame kernels as the runtime reports them
(xr_kernel_info_get): the occupancy the kernels ask for is the one they get."""
from __future__ import annotations

import pytest

from tests.helpers import say
from xvk_ray_amd import native as N

pytestmark = pytest.mark.gpu


def test_kernel_resources():
    info = {k["name"]: k for k in N.kernel_info()}
    say(f"[kernels] {info}")
    assert set(info) == {"k_initial_sample", "k_tail_bounce", "k_temporal", "k_spatial", "k_plain", "k_converge"}
    for k in info.values():
        assert 0 < k["vgprs"] <= 512 and k["blocks_per_cu"] >= 1 and k["block"] in (256,)
    # the spatial pass and the initial-sample head: 20 stack levels of 256
    # lanes in LDS, 7 waves per SIMD
    for name in ("k_spatial", "k_initial_sample"):
        assert info[name]["lds_bytes"] >= 20 * 256 * 4
        assert info[name]["waves_per_simd"] == 7
    # the other traversal kernels: 24 levels, 6 waves
    for name in ("k_tail_bounce", "k_plain"):
        assert info[name]["lds_bytes"] >= 24 * 256 * 4
        assert info[name]["waves_per_simd"] == 6
