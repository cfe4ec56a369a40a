"""MI355X (gfx950) path-tracing and ReSTIR GI layer for XVK_Ray's frame.

The compute lives in the native libraries under xvk_ray_amd/lib (HIP kernels
behind the C ABI of include/xr_api.h, and the C++ host driver behind
include/xr_host.h); this package only binds them (xvk_ray_amd.native).
"""
from .native import (App, Camera, Context, Scene, XrError, default_params, host, lib, make_ubo)  # noqa: F401

__all__ = ["App", "Camera", "Context", "Scene", "XrError", "default_params", "host", "lib", "make_ubo"]
