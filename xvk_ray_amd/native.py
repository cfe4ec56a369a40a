"""ctypes bindings of the two native libraries.

libxr.so       -- the HIP layer (include/xr_api.h)
libxr_host.so  -- the C++ host side: synthetic scenes, camera, frame driver (include/xr_host.h)

Loading fails loudly when a library is missing: there is no Python or CPU
fallback for any computation.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

XR_OK = 0
XR_ERR_INVALID_ARGUMENT = 1
XR_ERR_DEVICE = 2
XR_ERR_OUT_OF_MEMORY = 3
XR_ERR_STATE = 4
XR_ERR_UNSUPPORTED = 5

XR_MODE_PLAIN = 1
XR_MODE_RESTIR = 2
XR_VIS_REFERENCE = 0
XR_VIS_TRACE = 1
XR_PASS_INITIAL, XR_PASS_TEMPORAL, XR_PASS_SPATIAL = 1, 2, 4
XR_BUILDER_LBVH = 0
XR_BUILDER_PLOC = 1
XR_BUILDER_SAH = 2
XR_BUILDER_SAH_HOST = 3
XRH_GLTF_UNTEXTURED = 0x1

XR_BUF_ACCUMULATION = 1
XR_BUF_OUTPUT = 2
XR_BUF_SAMPLES = 9
XR_BUF_TEMPORAL = 10
XR_BUF_SPATIAL = 11
XR_BUF_GBUF_POS = 12
XR_BUF_GBUF_NRM = 13
XR_BUF_GBUF_MATUV = 14
XR_BUF_MOTION = 15
XR_BUF_RADIANCE = 16
XR_BUF_CONVERGED = 17
XR_FLAG_RADIANCE = 1
XR_FLAG_MLP_CACHE = 2
XR_MLP_INPUTS = 96
XR_MLP_PARAMS = 23043
XR_FLAG_MLP_FP8 = 4
XR_MLP_WEIGHTS_BF16 = 0
XR_MLP_WEIGHTS_FP8 = 1
XR_MLP_FP8_BYTES = 22528
XR_LAYOUT_REFERENCE = 0
XR_LAYOUT_NATIVE = 1

MATERIAL_LAMBERTIAN = 0
MATERIAL_DIFFUSE_LIGHT = 4


class XrVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("normal", C.c_float * 3), ("tex_coord", C.c_float * 2),
                ("material_index", C.c_int32)]


class XrMaterial(C.Structure):
    _fields_ = [("base_color", C.c_float * 4), ("base_color_texture_index", C.c_int32), ("fuzziness", C.c_float),
                ("refraction_index", C.c_float), ("material_model", C.c_uint32)]


class XrUbo(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("view_prev", C.c_float * 16), ("projection", C.c_float * 16),
                ("projection_prev", C.c_float * 16), ("view_inverse", C.c_float * 16),
                ("projection_inverse", C.c_float * 16), ("spp", C.c_uint32), ("total_number_of_samples", C.c_uint32),
                ("number_of_bounces", C.c_uint32), ("random_seed", C.c_uint32), ("has_sky_box", C.c_uint32),
                ("reuse", C.c_uint32), ("current_frame", C.c_uint32)]


class XrSceneDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32), ("indices", C.c_void_p),
                ("index_count", C.c_uint32), ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("offsets", C.c_void_p), ("instance_count", C.c_uint32)]


class XrTexture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba8", C.c_void_p)]


class XrRestirParams(C.Structure):
    _fields_ = [("temporal_m_cap", C.c_uint32), ("spatial_m_cap", C.c_uint32), ("spatial_samples", C.c_uint32),
                ("spatial_radius", C.c_float), ("visibility", C.c_uint32), ("plain_max_depth", C.c_uint32),
                ("passes", C.c_uint32), ("flags", C.c_uint32), ("mlp_spread", C.c_float), ("mlp_lr", C.c_float),
                ("mlp_train_pixels", C.c_uint32), ("mlp_loss_scale", C.c_float), ("reserved", C.c_uint32 * 4)]


class XrAccelInfo(C.Structure):
    _fields_ = [("triangle_count", C.c_uint32), ("node_count", C.c_uint32), ("max_depth", C.c_uint32),
                ("builder", C.c_uint32), ("build_ms", C.c_float), ("node_bytes", C.c_uint64),
                ("triangle_bytes", C.c_uint64)]


class XrKernelInfo(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("vgprs", C.c_uint32), ("scratch_bytes", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("block", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("waves_per_simd", C.c_uint32)]


class XrhSettings(C.Structure):
    _fields_ = [("rendering_mode", C.c_uint32), ("spp", C.c_uint32), ("number_of_bounces", C.c_uint32),
                ("max_spp", C.c_uint32), ("fov", C.c_float), ("enable_ray_accumulation", C.c_uint32),
                ("enable_sky_lighting", C.c_uint32), ("enable_reuse", C.c_uint32)]


assert C.sizeof(XrVertex) == 36 and C.sizeof(XrMaterial) == 32 and C.sizeof(XrUbo) == 412
assert C.sizeof(XrRestirParams) == 64

# numpy views of the reference AoS layouts (ReSTIRGICommon.h:8-32)
SAMPLE_DTYPE = np.dtype([("x_view", "<f4", 3), ("material_id", "<u4"), ("n_view", "<f4", 3), ("rand_seed", "<u4"),
                         ("x_sample", "<f4", 3), ("p_q", "<f4"), ("n_sample", "<f4", 3), ("pad1", "<u4"),
                         ("l_out", "<f4", 3), ("pad2", "<u4"), ("f", "<f4", 3), ("pad3", "<u4")])
RESERVOIR_DTYPE = np.dtype([("w_sum", "<f4"), ("w", "<f4"), ("m", "<u4"), ("pad", "<f4"), ("s", SAMPLE_DTYPE)])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("pad0", "<u4"), ("instance", "<u4"),
                      ("primitive", "<u4"), ("triangle", "<u4"), ("pad1", "<u4")])
VERTEX_DTYPE = np.dtype([("pos", "<f4", 3), ("normal", "<f4", 3), ("tex_coord", "<f4", 2), ("material_index", "<i4")])
MATERIAL_DTYPE = np.dtype([("base_color", "<f4", 4), ("base_color_texture_index", "<i4"), ("fuzziness", "<f4"),
                           ("refraction_index", "<f4"), ("material_model", "<u4")])
assert SAMPLE_DTYPE.itemsize == 96 and RESERVOIR_DTYPE.itemsize == 112
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32

XR_API_SYMBOLS = [
    "xr_default_params", "xr_create", "xr_build_accel", "xr_resize", "xr_render", "xr_sync", "xr_destroy",
    "xr_read_buffer", "xr_write_buffer", "xr_read_image", "xr_buffer_bytes", "xr_trace_rays",
    "xr_trace_rays_device", "xr_pass_times", "xr_pass_time_totals", "xr_reset_pass_times", "xr_counters",
    "xr_reset_counters", "xr_get_accel_info", "xr_kernel_count", "xr_kernel_info_get",
    "xr_read_accel", "xr_stream", "xr_converge", "xr_converge_reset", "xr_converged_spp", "xr_rmse",
    "xr_mlp_reset", "xr_mlp_eval", "xr_mlp_gradient", "xr_mlp_gradient_noalloc", "xr_mlp_train_rows", "xr_mlp_set_weight_format",
    "xr_mlp_fp8_weights", "xr_mlp_params",
    "xr_batch_create", "xr_batch_render", "xr_batch_frame_time", "xr_batch_set_graph", "xr_batch_config_count",
    "xr_batch_config_info", "xr_batch_config_persistent", "xr_batch_set_config", "xr_batch_config", "xr_batch_autotune", "xr_batch_reset",
    "xr_batch_counters", "xr_batch_reset_counters", "xr_batch_buffer_bytes", "xr_batch_read_buffer",
    "xr_batch_destroy",
    "xr_last_error", "xr_global_last_error", "xr_set_textures",
]
XR_HOST_SYMBOLS = [
    "xrh_scene_cornell", "xrh_scene_terrain", "xrh_scene_gltf", "xrh_scene_desc", "xrh_scene_textures",
    "xrh_scene_triangle_count", "xrh_scene_destroy", "xrh_png_decode",
    "xrh_camera_static", "xrh_camera_keyframes", "xrh_camera_orbit", "xrh_camera_view", "xrh_camera_destroy", "xrh_perspective",
    "xrh_inverse", "xrh_app_create", "xrh_app_draw_frame", "xrh_app_last_ubo", "xrh_app_ctx", "xrh_app_last_error",
    "xrh_app_destroy", "xrh_make_ubo", "xrh_png_encode", "xrh_write_png",
]

_lib = None
_host = None


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build the native libraries first (python -m xvk_ray_amd.build)")
    return C.CDLL(str(path))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = _load("libxr.so")
        P, U32, I = C.c_void_p, C.c_uint32, C.c_int
        L.xr_default_params.argtypes = [C.POINTER(XrRestirParams)]
        L.xr_default_params.restype = None
        L.xr_create.argtypes = [C.POINTER(XrSceneDesc), U32, U32, C.POINTER(P)]
        L.xr_build_accel.argtypes = [P, I]
        L.xr_resize.argtypes = [P, U32, U32]
        L.xr_render.argtypes = [P, C.POINTER(XrUbo), I, C.POINTER(XrRestirParams)]
        L.xr_sync.argtypes = [P]
        L.xr_destroy.argtypes = [P]
        L.xr_destroy.restype = None
        L.xr_read_buffer.argtypes = [P, I, P, C.c_size_t, I]
        L.xr_write_buffer.argtypes = [P, I, P, C.c_size_t, I]
        L.xr_read_image.argtypes = [P, P, C.c_size_t]
        L.xr_set_textures.argtypes = [P, P, U32]
        L.xr_buffer_bytes.argtypes = [P, I, I]
        L.xr_buffer_bytes.restype = C.c_size_t
        L.xr_trace_rays.argtypes = [P, P, U32, P, I]
        L.xr_trace_rays_device.argtypes = [P, P, U32, P, I]
        L.xr_pass_times.argtypes = [P, C.POINTER(C.c_float), U32]
        L.xr_converge.argtypes = [P, C.POINTER(XrUbo), U32, U32]
        L.xr_converge_reset.argtypes = [P]
        L.xr_converged_spp.argtypes = [P]
        L.xr_converged_spp.restype = C.c_uint64
        L.xr_rmse.argtypes = [P, C.POINTER(C.c_double)]
        L.xr_pass_time_totals.argtypes = [P, C.POINTER(C.c_double), U32, C.POINTER(U32)]
        L.xr_reset_pass_times.argtypes = [P]
        L.xr_counters.argtypes = [P, C.POINTER(C.c_uint64), U32]
        L.xr_reset_counters.argtypes = [P]
        L.xr_get_accel_info.argtypes = [P, C.POINTER(XrAccelInfo)]
        L.xr_kernel_count.argtypes = []
        L.xr_kernel_count.restype = U32
        L.xr_kernel_info_get.argtypes = [U32, C.POINTER(XrKernelInfo)]
        L.xr_read_accel.argtypes = [P, P, C.c_size_t, P, C.c_size_t]
        L.xr_stream.argtypes = [P]
        L.xr_stream.restype = P
        L.xr_mlp_reset.argtypes = [P, C.c_uint64]
        L.xr_mlp_eval.argtypes = [P, P, U32, P]
        L.xr_mlp_gradient.argtypes = [P, P, P, U32, P]
        L.xr_mlp_gradient_noalloc.argtypes = [P, P, P, U32, P]
        L.xr_mlp_train_rows.argtypes = [P, P, P, U32, C.c_float]
        L.xr_mlp_params.argtypes = [P, P, I]
        L.xr_mlp_set_weight_format.argtypes = [P, I]
        L.xr_mlp_fp8_weights.argtypes = [P, P, P]
        L.xr_batch_create.argtypes = [P, U32, P, P, C.POINTER(P)]
        L.xr_batch_render.argtypes = [P, P, C.POINTER(XrRestirParams)]
        L.xr_batch_frame_time.argtypes = [P, C.POINTER(C.c_float)]
        L.xr_batch_set_graph.argtypes = [P, I]
        L.xr_batch_config_count.argtypes = []
        L.xr_batch_config_count.restype = U32
        L.xr_batch_config_info.argtypes = [U32, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]
        L.xr_batch_config_persistent.argtypes = [U32, C.POINTER(C.c_int)]
        L.xr_batch_set_config.argtypes = [P, U32]
        L.xr_batch_config.argtypes = [P]
        L.xr_batch_config.restype = U32
        L.xr_batch_autotune.argtypes = [P, P, C.POINTER(XrRestirParams), U32, P]
        L.xr_batch_reset.argtypes = [P]
        L.xr_batch_counters.argtypes = [P, C.POINTER(C.c_uint64), U32]
        L.xr_batch_reset_counters.argtypes = [P]
        L.xr_batch_buffer_bytes.argtypes = [P, U32, I, I]
        L.xr_batch_buffer_bytes.restype = C.c_size_t
        L.xr_batch_read_buffer.argtypes = [P, U32, I, P, C.c_size_t, I]
        L.xr_batch_destroy.argtypes = [P]
        L.xr_batch_destroy.restype = None
        L.xr_last_error.argtypes = [P]
        L.xr_last_error.restype = C.c_char_p
        L.xr_global_last_error.argtypes = []
        L.xr_global_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def host() -> C.CDLL:
    global _host
    if _host is None:
        lib()  # libxr_host links against libxr
        H = _load("libxr_host.so")
        P, U32, U64, F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
        H.xrh_scene_cornell.argtypes = [U64, C.POINTER(P)]
        H.xrh_scene_terrain.argtypes = [U64, U32, U32, U32, C.POINTER(P)]
        H.xrh_scene_gltf.argtypes = [C.POINTER(C.c_char_p), U32, U32, C.POINTER(P), C.c_char_p, C.c_size_t]
        H.xrh_png_encode.argtypes = [P, U32, U32, P, C.c_size_t, C.POINTER(C.c_size_t)]
        H.xrh_png_decode.argtypes = [P, C.c_size_t, C.POINTER(U32), C.POINTER(U32), P, C.c_size_t]
        H.xrh_scene_textures.argtypes = [P, C.POINTER(C.POINTER(XrTexture)), C.POINTER(U32)]
        H.xrh_write_png.argtypes = [C.c_char_p, P, U32, U32]
        H.xrh_scene_desc.argtypes = [P, C.POINTER(XrSceneDesc)]
        H.xrh_scene_triangle_count.argtypes = [P]
        H.xrh_scene_triangle_count.restype = U32
        H.xrh_scene_destroy.argtypes = [P]
        H.xrh_scene_destroy.restype = None
        H.xrh_camera_static.argtypes = [C.POINTER(F), C.POINTER(F), C.POINTER(P)]
        H.xrh_camera_orbit.argtypes = [U64, U32, U32, C.POINTER(P)]
        H.xrh_camera_keyframes.argtypes = [C.POINTER(F), C.POINTER(F), U32, C.POINTER(P)]
        H.xrh_camera_view.argtypes = [P, U32, C.POINTER(F)]
        H.xrh_camera_destroy.argtypes = [P]
        H.xrh_camera_destroy.restype = None
        H.xrh_perspective.argtypes = [F, F, F, F, C.POINTER(F)]
        H.xrh_perspective.restype = None
        H.xrh_inverse.argtypes = [C.POINTER(F), C.POINTER(F)]
        H.xrh_app_create.argtypes = [C.POINTER(XrhSettings), P, P, U32, U32, C.POINTER(P)]
        H.xrh_app_draw_frame.argtypes = [P, C.POINTER(XrRestirParams)]
        H.xrh_app_last_ubo.argtypes = [P, C.POINTER(XrUbo)]
        H.xrh_app_ctx.argtypes = [P]
        H.xrh_app_ctx.restype = P
        H.xrh_app_last_error.argtypes = [P]
        H.xrh_app_last_error.restype = C.c_char_p
        H.xrh_app_destroy.argtypes = [P]
        H.xrh_app_destroy.restype = None
        H.xrh_make_ubo.argtypes = [C.POINTER(F), C.POINTER(F), F, U32, U32, U32, U32, U32, U32, U32, U32,
                                   C.POINTER(XrUbo)]
        H.xrh_make_ubo.restype = None
        _host = H
    return _host


class XrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"xr error {code}: {msg}")
        self.code = code


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def kernel_info() -> list[dict]:
    """Resources of the frame kernels (xr_kernel_info_get); needs a device."""
    L = lib()
    out = []
    for i in range(L.xr_kernel_count()):
        k = XrKernelInfo()
        rc = L.xr_kernel_info_get(i, C.byref(k))
        if rc:
            raise XrError(rc, L.xr_global_last_error().decode())
        out.append({"name": k.name.decode(), "vgprs": k.vgprs, "scratch_bytes": k.scratch_bytes,
                    "lds_bytes": k.lds_bytes, "block": k.block, "blocks_per_cu": k.blocks_per_cu,
                    "waves_per_simd": k.waves_per_simd})
    return out


def default_params(**overrides) -> XrRestirParams:
    p = XrRestirParams()
    lib().xr_default_params(C.byref(p))
    names = {f[0] for f in XrRestirParams._fields_}
    for k, v in overrides.items():
        if k not in names:
            raise AttributeError(f"xr_restir_params has no field {k!r}")
        setattr(p, k, v)
    return p


def ubo_to_bytes(u: XrUbo) -> bytes:
    return bytes(memoryview(u))


class Scene:
    """A synthetic scene built by the C++ host (xrh_scene_*), or one given as arrays."""

    def __init__(self, handle=None, arrays=None):
        self._h = handle
        self._arrays = arrays  # keeps numpy arrays alive for array-made scenes

    @classmethod
    def cornell(cls, seed: int = 1) -> "Scene":
        h = C.c_void_p()
        rc = host().xrh_scene_cornell(seed, C.byref(h))
        if rc:
            raise XrError(rc, "xrh_scene_cornell")
        return cls(handle=h)

    @classmethod
    def terrain(cls, seed: int = 1, grid: int = 1024, spheres: int = 512, lights: int = 64) -> "Scene":
        h = C.c_void_p()
        rc = host().xrh_scene_terrain(seed, grid, spheres, lights, C.byref(h))
        if rc:
            raise XrError(rc, "xrh_scene_terrain")
        return cls(handle=h)

    @classmethod
    def gltf(cls, *paths, untextured: bool = False) -> "Scene":
        """glTF 2.0 files, one instance each, loaded as the reference's
        AglTFModel / AScene do (see xrh_scene_gltf in include/xr_host.h)."""
        if not paths:
            raise ValueError("at least one path")
        arr = (C.c_char_p * len(paths))(*[os.fsencode(os.fspath(p)) for p in paths])
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = host().xrh_scene_gltf(arr, len(paths), XRH_GLTF_UNTEXTURED if untextured else 0, C.byref(h), err,
                                   len(err))
        if rc:
            raise XrError(rc, "xrh_scene_gltf: " + err.value.decode(errors="replace"))
        return cls(handle=h)

    @classmethod
    def from_arrays(cls, vertices: np.ndarray, indices: np.ndarray, materials: np.ndarray,
                    offsets: np.ndarray, textures=()) -> "Scene":
        """textures: [(width, height, RGBA8 uint8 array)], given to xr_set_textures by Context."""
        v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        m = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        o = np.ascontiguousarray(offsets, dtype=np.uint32).reshape(-1)
        t = [(int(w), int(h), np.ascontiguousarray(px, np.uint8).reshape(-1)) for w, h, px in textures]
        return cls(arrays=(v, i, m, o, t))

    def desc(self) -> XrSceneDesc:
        d = XrSceneDesc()
        if self._h is not None:
            rc = host().xrh_scene_desc(self._h, C.byref(d))
            if rc:
                raise XrError(rc, "xrh_scene_desc")
        else:
            v, i, m, o = self._arrays[:4]
            d.vertices, d.vertex_count = v.ctypes.data, len(v)
            d.indices, d.index_count = i.ctypes.data, len(i)
            d.materials, d.material_count = m.ctypes.data, len(m)
            d.offsets, d.instance_count = o.ctypes.data, len(o) // 2
        return d

    def arrays(self):
        """(vertices, indices, materials, offsets) as numpy views/copies."""
        if self._arrays is not None:
            return self._arrays[:4]
        d = self.desc()

        def view(ptr, n, ctype, dtype):
            if n == 0 or not ptr:  # an empty std::vector has no storage
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array((ctype * n).from_address(ptr)).view(dtype).copy()

        v = view(d.vertices, 36 * d.vertex_count, C.c_uint8, VERTEX_DTYPE)
        i = view(d.indices, d.index_count, C.c_uint32, np.uint32)
        m = view(d.materials, 32 * d.material_count, C.c_uint8, MATERIAL_DTYPE)
        o = view(d.offsets, 2 * d.instance_count, C.c_uint32, np.uint32)
        return v, i, m, o

    def textures(self) -> list:
        """[(width, height, RGBA8 uint8 array)] of a loaded glTF scene (empty otherwise)."""
        if self._h is None:
            return list(self._arrays[4]) if len(self._arrays) > 4 else []
        ptr = C.POINTER(XrTexture)()
        n = C.c_uint32()
        rc = host().xrh_scene_textures(self._h, C.byref(ptr), C.byref(n))
        if rc:
            raise XrError(rc, "xrh_scene_textures")
        out = []
        for k in range(n.value):
            t = ptr[k]
            nb = 4 * t.width * t.height
            px = np.ctypeslib.as_array((C.c_uint8 * nb).from_address(t.rgba8)).copy()
            out.append((t.width, t.height, px))
        return out

    @property
    def triangle_count(self) -> int:
        return self.desc().index_count // 3

    def close(self):
        if self._h is not None:
            host().xrh_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Camera:
    def __init__(self, handle):
        self._h = handle

    @classmethod
    def static(cls, eye, target) -> "Camera":
        e = np.asarray(eye, np.float32)
        t = np.asarray(target, np.float32)
        h = C.c_void_p()
        rc = host().xrh_camera_static(_fptr(e), _fptr(t), C.byref(h))
        if rc:
            raise XrError(rc, "xrh_camera_static")
        return cls(h)

    @classmethod
    def keyframes(cls, eyes, targets) -> "Camera":
        e = np.ascontiguousarray(eyes, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        assert e.shape == t.shape
        h = C.c_void_p()
        rc = host().xrh_camera_keyframes(_fptr(e), _fptr(t), len(e), C.byref(h))
        if rc:
            raise XrError(rc, "xrh_camera_keyframes")
        return cls(h)

    @classmethod
    def orbit(cls, seed: int = 2, n_ctrl: int = 8, frames: int = 128) -> "Camera":
        h = C.c_void_p()
        rc = host().xrh_camera_orbit(seed, n_ctrl, frames, C.byref(h))
        if rc:
            raise XrError(rc, "xrh_camera_orbit")
        return cls(h)

    def view(self, k: int) -> np.ndarray:
        m = np.zeros(16, np.float32)
        host().xrh_camera_view(self._h, k, _fptr(m))
        return m

    def close(self):
        if self._h is not None:
            host().xrh_camera_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_ubo(view, view_prev, fov, width, height, spp=1, total=1, bounces=4, sky=0, reuse=1, frame=1) -> XrUbo:
    u = XrUbo()
    v = np.ascontiguousarray(view, np.float32)
    vp = np.ascontiguousarray(view_prev, np.float32)
    host().xrh_make_ubo(_fptr(v), _fptr(vp), fov, width, height, spp, total, bounces, sky, reuse, frame, C.byref(u))
    return u


def _typed(out: np.ndarray, which: int, layout: int) -> np.ndarray:
    """View of raw buffer bytes with the element type of buffer `which`."""
    if layout == XR_LAYOUT_REFERENCE:
        if which == XR_BUF_SAMPLES:
            return out.view(SAMPLE_DTYPE)
        if which in (XR_BUF_TEMPORAL, XR_BUF_SPATIAL):
            return out.view(RESERVOIR_DTYPE)
    if which in (XR_BUF_OUTPUT, XR_BUF_MOTION):
        return out.view(np.uint32)
    if which in (XR_BUF_ACCUMULATION, XR_BUF_GBUF_POS, XR_BUF_GBUF_NRM, XR_BUF_GBUF_MATUV, XR_BUF_RADIANCE):
        return out.view(np.float32).reshape(-1, 4)
    return out.view(np.float32)


# Contexts still open when the interpreter exits (e.g. held by a traceback)
# are closed before the HIP runtime's own teardown, batches first.
_LIVE_CONTEXTS: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _close_live_contexts() -> None:
    for ctx in list(_LIVE_CONTEXTS):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """One xr_ctx: scene on the device, frame buffers, acceleration structure."""

    def __init__(self, scene: Scene, width: int, height: int, build: bool = True, builder: int = XR_BUILDER_LBVH):
        self.scene = scene
        d = scene.desc()
        h = C.c_void_p()
        rc = lib().xr_create(C.byref(d), width, height, C.byref(h))
        if rc:
            raise XrError(rc, lib().xr_global_last_error().decode())
        self._h = h
        # open batches, held until they close: a batch must not outlive its
        # context, also when both are collected together (weak references
        # would already be cleared when the context's finalizer runs)
        self._batches: dict[int, "Batch"] = {}
        _LIVE_CONTEXTS.add(self)
        self.width, self.height = width, height
        tex = scene.textures()
        if tex:
            self.set_textures(tex)
        if build:
            self.build_accel(builder)

    def set_textures(self, textures) -> None:
        """[(width, height, RGBA8 uint8 array)] -> xr_set_textures."""
        self._tex_keep = [np.ascontiguousarray(px, np.uint8) for _, _, px in textures]
        arr = (XrTexture * len(textures))()
        for k, ((w, h, _), px) in enumerate(zip(textures, self._tex_keep)):
            arr[k].width, arr[k].height, arr[k].rgba8 = w, h, px.ctypes.data
        self.check(lib().xr_set_textures(self._h, arr, len(textures)))

    def build_accel(self, builder: int = XR_BUILDER_LBVH) -> None:
        self.check(lib().xr_build_accel(self._h, builder))

    def check(self, rc: int) -> None:
        if rc:
            raise XrError(rc, lib().xr_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def render(self, ubo: XrUbo, mode: int = XR_MODE_RESTIR, params: XrRestirParams | None = None) -> None:
        self.check(lib().xr_render(self._h, C.byref(ubo), mode, C.byref(params) if params is not None else None))

    def sync(self) -> None:
        self.check(lib().xr_sync(self._h))

    def resize(self, w: int, h: int) -> None:
        self.check(lib().xr_resize(self._h, w, h))
        self.width, self.height = w, h

    def buffer_bytes(self, which: int, layout: int) -> int:
        return lib().xr_buffer_bytes(self._h, which, layout)

    def read(self, which: int, layout: int = XR_LAYOUT_REFERENCE) -> np.ndarray:
        n = self.buffer_bytes(which, layout)
        out = np.empty(n, np.uint8)
        self.check(lib().xr_read_buffer(self._h, which, out.ctypes.data, n, layout))
        return _typed(out, which, layout)

    def read_image(self) -> np.ndarray:
        """The tone-mapped B8G8R8A8 output image (xr_read_image), one uint32 per pixel."""
        n = self.buffer_bytes(XR_BUF_OUTPUT, XR_LAYOUT_REFERENCE)
        out = np.empty(n // 4, np.uint32)
        self.check(lib().xr_read_image(self._h, out.ctypes.data, out.nbytes))
        return out

    def save_png(self, path) -> None:
        """The output image as an 8-bit RGB PNG (xrh_write_png)."""
        img = self.read_image()
        rc = host().xrh_write_png(os.fsencode(os.fspath(path)), img.ctypes.data, self.width, self.height)
        if rc:
            raise XrError(rc, f"xrh_write_png {path}")

    def write(self, which: int, data: np.ndarray, layout: int = XR_LAYOUT_REFERENCE) -> None:
        a = np.ascontiguousarray(data)
        self.check(lib().xr_write_buffer(self._h, which, a.ctypes.data, a.nbytes, layout))

    def trace(self, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        r = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(r), HIT_DTYPE)
        self.check(lib().xr_trace_rays(self._h, r.ctypes.data, len(r), hits.ctypes.data, int(any_hit)))
        return hits

    def pass_times(self) -> list[float]:
        ms = (C.c_float * 4)()
        n = lib().xr_pass_times(self._h, ms, 4)
        if n < 0:
            self.check(n)
        return [ms[i] for i in range(n)]

    def converge(self, ubo: XrUbo, spp: int, frame_base: int) -> None:
        self.check(lib().xr_converge(self._h, C.byref(ubo), spp, frame_base))

    def converge_reset(self) -> None:
        self.check(lib().xr_converge_reset(self._h))

    def converged_spp(self) -> int:
        return int(lib().xr_converged_spp(self._h))

    def converged_mean(self) -> np.ndarray:
        """(n, 3) float64 mean of the converged image."""
        n = self.buffer_bytes(XR_BUF_CONVERGED, XR_LAYOUT_NATIVE)
        out = np.empty(n // 8, np.float64)
        self.check(lib().xr_read_buffer(self._h, XR_BUF_CONVERGED, out.ctypes.data, n, XR_LAYOUT_NATIVE))
        return out.reshape(-1, 3) / max(1, self.converged_spp())

    def rmse(self) -> float:
        r = C.c_double()
        self.check(lib().xr_rmse(self._h, C.byref(r)))
        return float(r.value)

    # ---- MLP radiance cache ----
    def mlp_reset(self, seed: int = 1) -> None:
        self.check(lib().xr_mlp_reset(self._h, seed))

    def mlp_eval(self, inputs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, XR_MLP_INPUTS)
        out = np.zeros((x.shape[0], 3), np.float32)
        self.check(lib().xr_mlp_eval(self._h, x.ctypes.data, x.shape[0], out.ctypes.data))
        return out

    def mlp_gradient(self, inputs: np.ndarray, targets: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, XR_MLP_INPUTS)
        y = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        assert x.shape[0] == y.shape[0]
        g = np.zeros(XR_MLP_PARAMS, np.float32)
        self.check(lib().xr_mlp_gradient(self._h, x.ctypes.data, y.ctypes.data, x.shape[0], g.ctypes.data))
        return g

    def mlp_train_rows(self, inputs: np.ndarray, targets: np.ndarray, lr: float = 1e-3) -> None:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, XR_MLP_INPUTS)
        y = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        assert x.shape[0] == y.shape[0]
        self.check(lib().xr_mlp_train_rows(self._h, x.ctypes.data, y.ctypes.data, x.shape[0], lr))

    def mlp_set_weight_format(self, fmt: int) -> None:
        self.check(lib().xr_mlp_set_weight_format(self._h, fmt))

    def mlp_fp8_weights(self) -> tuple[np.ndarray, np.ndarray]:
        """(E4M3 codes of the hidden layers' W^T, per-layer scales)."""
        codes = np.zeros(XR_MLP_FP8_BYTES, np.uint8)
        scales = np.zeros(5, np.float32)
        self.check(lib().xr_mlp_fp8_weights(self._h, codes.ctypes.data, scales.ctypes.data))
        return codes, scales

    def mlp_params(self, params: np.ndarray | None = None) -> np.ndarray:
        """FP32 master parameters; writes `params` into the network first when given."""
        buf = np.zeros(XR_MLP_PARAMS, np.float32)
        if params is not None:
            buf[:] = np.asarray(params, np.float32).reshape(-1)
        self.check(lib().xr_mlp_params(self._h, buf.ctypes.data, 0 if params is None else 1))
        return buf

    def pass_time_totals(self) -> tuple[list[float], int]:
        """Summed per-pass milliseconds (initial or plain, temporal, spatial, cache) and frames."""
        ms = (C.c_double * 4)()
        frames = C.c_uint32()
        self.check(lib().xr_pass_time_totals(self._h, ms, 4, C.byref(frames)))
        return [ms[i] for i in range(4)], int(frames.value)

    def reset_pass_times(self) -> None:
        self.check(lib().xr_reset_pass_times(self._h))

    def counters(self) -> tuple[int, int]:
        """(extend rays, shadow rays) since the last reset."""
        return self.counters3()[:2]

    def counters3(self) -> tuple[int, int, int]:
        """(extend rays, shadow rays, valid primary hits) since the last reset."""
        return self.counters_all()[:3]

    def counters_all(self) -> tuple[int, int, int, int, int]:
        """(extend rays, shadow rays, valid primary hits, cache queries, cache training samples)."""
        c = (C.c_uint64 * 5)()
        self.check(lib().xr_counters(self._h, c, 5))
        return tuple(int(c[i]) for i in range(5))

    def reset_counters(self) -> None:
        self.check(lib().xr_reset_counters(self._h))

    def accel_info(self) -> XrAccelInfo:
        info = XrAccelInfo()
        self.check(lib().xr_get_accel_info(self._h, C.byref(info)))
        return info

    def read_accel(self):
        """(nodes, triangles): nodes as (n, 16) float32 in the layout xr_read_accel documents (grid
        origin in columns 0-2, exponents and child count in the bits of column 3, plane bytes in
        columns 4-9, links in columns 12-15); triangles as (m, 12) float32 (w columns hold id /
        instance / primitive bits)."""
        info = self.accel_info()
        nodes = np.zeros((info.node_count, 16), np.float32)
        tris = np.zeros((info.triangle_count, 12), np.float32)
        self.check(lib().xr_read_accel(self._h, nodes.ctypes.data, nodes.nbytes, tris.ctypes.data, tris.nbytes))
        return nodes, tris

    def close(self):
        if getattr(self, "_h", None) is not None:
            for b in list(getattr(self, "_batches", {}).values()):
                b.close()
            lib().xr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class App:
    """The C++ frame driver (xrh_app): PathTracer::DrawFrame without a window."""

    def __init__(self, scene: Scene, camera: Camera, width: int, height: int, mode: int = XR_MODE_RESTIR,
                 spp: int = 1, bounces: int = 4, max_spp: int = 1000, fov: float = 45.0, accumulate: bool = False,
                 sky: bool = False, reuse: bool = True):
        self.scene, self.camera = scene, camera
        s = XrhSettings(mode, spp, bounces, max_spp, fov, int(accumulate), int(sky), int(reuse))
        h = C.c_void_p()
        rc = host().xrh_app_create(C.byref(s), scene._h if scene._h is not None else None, camera._h, width, height,
                                   C.byref(h))
        if rc:
            raise XrError(rc, lib().xr_global_last_error().decode())
        self._h = h
        self.width, self.height = width, height

    @property
    def ctx_handle(self):
        return host().xrh_app_ctx(self._h)

    def draw_frame(self, params: XrRestirParams | None = None) -> None:
        rc = host().xrh_app_draw_frame(self._h, C.byref(params) if params is not None else None)
        if rc:
            raise XrError(rc, host().xrh_app_last_error(self._h).decode())

    def last_ubo(self) -> XrUbo:
        u = XrUbo()
        host().xrh_app_last_ubo(self._h, C.byref(u))
        return u

    def context(self) -> "BorrowedContext":
        return BorrowedContext(self.ctx_handle, self.width, self.height)

    def close(self):
        if getattr(self, "_h", None) is not None:
            host().xrh_app_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BorrowedContext(Context):
    """A Context view of a context owned by an App (not destroyed by this object)."""

    def __init__(self, handle, width, height):
        self._h = handle
        self.width, self.height = width, height
        self.scene = None

    def close(self):
        self._h = None


class Batch:
    """Views rendered together on the scene of a Context (xr_batch)."""

    def __init__(self, ctx: "Context", sizes):
        self.ctx = ctx
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        ws = np.array([w for w, _ in self.sizes], np.uint32)
        hs = np.array([h for _, h in self.sizes], np.uint32)
        h = C.c_void_p()
        ctx.check(lib().xr_batch_create(ctx.handle, len(self.sizes), ws.ctypes.data, hs.ctypes.data, C.byref(h)))
        self._h = h
        ctx._batches[id(self)] = self

    @staticmethod
    def config_count() -> int:
        return int(lib().xr_batch_config_count())

    @staticmethod
    def config_info(k: int) -> tuple[int, int, int]:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        if lib().xr_batch_config_info(k, C.byref(a), C.byref(b), C.byref(c)):
            raise XrError(XR_ERR_INVALID_ARGUMENT, lib().xr_global_last_error().decode())
        return int(a.value), int(b.value), int(c.value)

    @staticmethod
    def config_persistent(k: int) -> bool:
        p = C.c_int()
        if lib().xr_batch_config_persistent(k, C.byref(p)):
            raise XrError(XR_ERR_INVALID_ARGUMENT, lib().xr_global_last_error().decode())
        return p.value == 1

    def _ubos(self, ubos):
        assert len(ubos) == len(self.sizes)
        return (XrUbo * len(ubos))(*ubos)

    def render(self, ubos, params: XrRestirParams | None = None) -> None:
        arr = self._ubos(ubos)
        self.ctx.check(lib().xr_batch_render(self._h, arr, C.byref(params) if params is not None else None))

    def frame_time(self) -> float:
        ms = C.c_float()
        self.ctx.check(lib().xr_batch_frame_time(self._h, C.byref(ms)))
        return float(ms.value)

    def set_graph(self, enable: bool) -> None:
        self.ctx.check(lib().xr_batch_set_graph(self._h, int(enable)))

    def set_config(self, k: int) -> None:
        self.ctx.check(lib().xr_batch_set_config(self._h, k))

    @property
    def config(self) -> int:
        return int(lib().xr_batch_config(self._h))

    def autotune(self, ubos, params: XrRestirParams | None = None, frames: int = 4) -> list[float]:
        arr = self._ubos(ubos)
        ms = (C.c_float * self.config_count())()
        self.ctx.check(lib().xr_batch_autotune(self._h, arr, C.byref(params) if params is not None else None, frames,
                                               ms))
        return [float(x) for x in ms]

    def reset(self) -> None:
        self.ctx.check(lib().xr_batch_reset(self._h))

    def counters(self) -> tuple[int, ...]:
        c = (C.c_uint64 * 5)()
        self.ctx.check(lib().xr_batch_counters(self._h, c, 5))
        return tuple(int(c[i]) for i in range(5))

    def reset_counters(self) -> None:
        self.ctx.check(lib().xr_batch_reset_counters(self._h))

    def read(self, view: int, which: int, layout: int = XR_LAYOUT_REFERENCE) -> np.ndarray:
        n = int(lib().xr_batch_buffer_bytes(self._h, view, which, layout))
        out = np.empty(n, np.uint8)
        self.ctx.check(lib().xr_batch_read_buffer(self._h, view, which, out.ctypes.data, n, layout))
        return _typed(out, which, layout)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().xr_batch_destroy(self._h)
            self._h = None
            self.ctx._batches.pop(id(self), None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

