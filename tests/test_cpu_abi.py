"""The C-ABI libraries load and export every function include/*.h declares,
and argument validation fails with the documented status codes before any
device work (no GPU needed)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_available
from xvk_ray_amd import native as N

INCLUDE = Path(__file__).resolve().parents[1] / "include"


def declared_functions(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?[\s\*]((?:xr|xrh)_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_expected_entry_points():
    api = declared_functions(INCLUDE / "xr_api.h")
    host = declared_functions(INCLUDE / "xr_host.h")
    assert "xr_create" in api and "xr_render" in api and "xr_trace_rays" in api
    assert "xrh_app_draw_frame" in host
    assert sorted(api) == sorted(N.XR_API_SYMBOLS), set(api) ^ set(N.XR_API_SYMBOLS)
    assert sorted(host) == sorted(N.XR_HOST_SYMBOLS), set(host) ^ set(N.XR_HOST_SYMBOLS)


@pytest.mark.parametrize("header,loader", [("xr_api.h", N.lib), ("xr_host.h", N.host)])
def test_library_exports_every_declared_symbol(header, loader):
    lib = loader()
    for name in declared_functions(INCLUDE / header):
        assert hasattr(lib, name), f"{name} declared in {header} but not exported"


def _scene_arrays():
    verts = np.zeros(3, N.VERTEX_DTYPE)
    verts["pos"] = [[0, 0, 1], [1, 0, 1], [0, 1, 1]]
    mats = np.zeros(1, N.MATERIAL_DTYPE)
    mats["base_color_texture_index"] = -1
    return verts, np.arange(3, dtype=np.uint32), mats, np.array([0, 0], np.uint32)


def _create(v, i, m, o, w=8, h=8):
    s = N.Scene.from_arrays(v, i, m, o)
    d = s.desc()
    h_ = C.c_void_p()
    rc = N.lib().xr_create(C.byref(d), w, h, C.byref(h_))
    return rc, h_, N.lib().xr_global_last_error().decode()


def test_create_rejects_bad_arguments_before_device_work():
    v, i, m, o = _scene_arrays()
    h = C.c_void_p()
    assert N.lib().xr_create(None, 8, 8, C.byref(h)) == N.XR_ERR_INVALID_ARGUMENT
    assert not h.value
    rc, h, msg = _create(v, i, m, o, w=0)
    assert rc == N.XR_ERR_INVALID_ARGUMENT and "resolution" in msg and not h.value
    rc, _, msg = _create(v, np.array([0, 1, 3], np.uint32), m, o)
    assert rc == N.XR_ERR_INVALID_ARGUMENT and "out of range" in msg
    rc, _, msg = _create(v, i[:2], m, o)
    assert rc == N.XR_ERR_INVALID_ARGUMENT
    m2 = m.copy()
    m2["base_color_texture_index"] = -2
    rc, _, msg = _create(v, i, m2, o)
    assert rc == N.XR_ERR_INVALID_ARGUMENT and "texture" in msg
    v2 = v.copy()
    v2["material_index"] = 5
    rc, _, msg = _create(v2, i, m, o)
    assert rc == N.XR_ERR_INVALID_ARGUMENT and "material" in msg
    rc, _, msg = _create(v, i, m, np.array([3, 0], np.uint32))
    assert rc == N.XR_ERR_INVALID_ARGUMENT
    # coordinates must be finite and within +-2^40 (the limit itself passes
    # validation and fails later only for want of a device)
    for bad in (np.nan, np.inf, -np.inf, 2.0**40 * (1 + 2.0**-23), -2.0**41):
        v3 = v.copy()
        v3["pos"][1, 2] = bad
        rc, h, msg = _create(v3, i, m, o)
        assert rc == N.XR_ERR_UNSUPPORTED and "2^40" in msg and not h.value
    v4 = v.copy()
    v4["pos"][1] = [2.0**40, -(2.0**40), 0.0]
    rc, h, msg = _create(v4, i, m, o)
    assert rc != N.XR_ERR_UNSUPPORTED
    if h.value:
        N.lib().xr_destroy(h)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_reports_device_error():
    v, i, m, o = _scene_arrays()
    rc, h, msg = _create(v, i, m, o)
    assert rc == N.XR_ERR_DEVICE and not h.value and msg


def test_kernel_info_arguments():
    L = N.lib()
    assert L.xr_kernel_count() == 6
    k = N.XrKernelInfo()
    assert L.xr_kernel_info_get(99, C.byref(k)) == N.XR_ERR_INVALID_ARGUMENT
    assert L.xr_kernel_info_get(0, None) == N.XR_ERR_INVALID_ARGUMENT


def test_batch_configurations():
    """Nine tile shape x stack depth combinations, each once with one
    workgroup per tile and once persistent, in that order."""
    L = N.lib()
    n = N.Batch.config_count()
    assert n == 18
    info = [N.Batch.config_info(k) for k in range(n)]
    assert info[:9] == info[9:]
    assert len(set(info[:9])) == 9
    for tw, th, lds in info:
        assert tw * th == 256 and lds in (16, 20, 24)
    assert [N.Batch.config_persistent(k) for k in range(n)] == [False] * 9 + [True] * 9
    p = C.c_int()
    assert L.xr_batch_config_persistent(n, C.byref(p)) == N.XR_ERR_INVALID_ARGUMENT
    assert L.xr_batch_config_persistent(0, None) == N.XR_ERR_INVALID_ARGUMENT


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_kernel_info_without_device():
    k = N.XrKernelInfo()
    assert N.lib().xr_kernel_info_get(0, C.byref(k)) == N.XR_ERR_DEVICE


def test_null_context_calls_are_safe():
    L = N.lib()
    L.xr_destroy(None)
    assert L.xr_sync(None) == N.XR_ERR_INVALID_ARGUMENT
    assert L.xr_render(None, None, 2, None) == N.XR_ERR_INVALID_ARGUMENT
    assert L.xr_build_accel(None, 0) == N.XR_ERR_INVALID_ARGUMENT
    assert L.xr_buffer_bytes(None, N.XR_BUF_SAMPLES, 0) == 0
    p = N.default_params()
    assert (p.temporal_m_cap, p.spatial_m_cap, p.spatial_samples, p.spatial_radius) == (30, 300, 6, 5.0)
    assert p.visibility == N.XR_VIS_REFERENCE and p.plain_max_depth == 0


def test_host_math_closed_forms():
    H = N.host()
    m = np.zeros(16, np.float32)
    H.xrh_perspective(90.0, 2.0, 0.001, 1000.0, N._fptr(m))
    M = m.reshape(4, 4)  # M[col][row]
    assert np.isclose(M[0, 0], 0.5) and np.isclose(M[1, 1], -1.0)  # 1/(aspect*tan45), -1/tan45
    assert M[2, 3] == -1.0 and np.isclose(M[2, 2], 1000.0 / (0.001 - 1000.0))
    inv = np.zeros(16, np.float32)
    assert H.xrh_inverse(N._fptr(m), N._fptr(inv)) == 1
    prod = inv.reshape(4, 4).T.astype(np.float64) @ m.reshape(4, 4).T.astype(np.float64)
    assert np.allclose(prod, np.eye(4), atol=1e-4)
    # the centre ray of a symmetric camera is the view's forward axis
    cam = N.Camera.static((0.5, 0.5, -1.2), (0.5, 0.5, 0.5))
    view = cam.view(0).reshape(4, 4).T.astype(np.float64)
    fwd = -view[2, :3]
    assert np.allclose(fwd, [0, 0, 1], atol=1e-6)
    assert np.allclose(view @ [0.5, 0.5, -1.2, 1.0], [0, 0, 0, 1], atol=1e-6)
    u = N.make_ubo(cam.view(0), cam.view(0), 60.0, 64, 64)
    pinv = np.frombuffer(bytes(u.projection_inverse), np.float32).reshape(4, 4).T
    t = pinv @ np.array([0, 0, 1, 1], np.float32)
    d = t[:3] / np.linalg.norm(t[:3])
    assert np.allclose(d, [0, 0, -1], atol=1e-6)  # NDC centre -> view-space -z


def test_synthetic_scene_sizes():
    c = N.Scene.cornell(1)
    assert c.triangle_count == 24 + 2 + 1280
    t = N.Scene.terrain(1, 64, 3, 2)
    assert t.triangle_count == 2 * 63 * 63 + 3 * 1280 + 2 * 2
    v, i, m, o = t.arrays()
    assert len(o) == 2 * (1 + 3 + 1)
    lights = m["material_model"] == N.MATERIAL_DIFFUSE_LIGHT
    assert lights.sum() == 2
    assert np.all((m["base_color"][lights][:, 0] * 4 >= 5) & (m["base_color"][lights][:, 0] * 4 <= 50))
    # same seed, same scene
    v2, i2, _, _ = N.Scene.terrain(1, 64, 3, 2).arrays()
    assert v.tobytes() == v2.tobytes() and i.tobytes() == i2.tobytes()
