"""ctypes binding of oracle/liboracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

P, U32, U64, F, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int


def build_oracle() -> None:
    """Compile oracle.cpp with its Makefile (gcc, no GPU)."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


class Oracle:
    _lib = None

    @classmethod
    def lib(cls) -> C.CDLL:
        if cls._lib is None:
            if not LIB.exists():
                build_oracle()
            L = C.CDLL(str(LIB))
            L.orc_num_threads.restype = I
            L.orc_set_threads.argtypes = [I]
            L.orc_set_threads.restype = None
            L.orc_init_seed.argtypes = [U32, U32, U32, U32]
            L.orc_init_seed.restype = U32
            L.orc_random_stream.argtypes = [U32, U32, P, P]
            L.orc_random_stream.restype = None
            L.orc_sincos.argtypes = [P, U32, P, P]
            L.orc_sincos.restype = None
            L.orc_onb.argtypes = [P, P, P]
            L.orc_onb.restype = None
            L.orc_sample_lambert.argtypes = [P, P, U32, P, P, P, P, P]
            L.orc_sample_lambert.restype = None
            L.orc_resolve.argtypes = [P, F]
            L.orc_resolve.restype = U32
            L.orc_aces.argtypes = [F]
            L.orc_aces.restype = F
            L.orc_to_half.argtypes = [F]
            L.orc_to_half.restype = C.c_uint16
            L.orc_from_half.argtypes = [C.c_uint16]
            L.orc_from_half.restype = F
            L.orc_update_reservoir.argtypes = [P, P, F, C.POINTER(U32)]
            L.orc_update_reservoir.restype = None
            L.orc_merge_reservoir.argtypes = [P, P, F, C.POINTER(U32)]
            L.orc_merge_reservoir.restype = None
            L.orc_scene_create.argtypes = [P, U32, P, U32, P, U32, P, U32]
            L.orc_scene_create.restype = P
            L.orc_scene_set_textures.argtypes = [P, P, U32, P, C.c_size_t]
            L.orc_scene_set_textures.restype = C.c_int
            L.orc_texture_sample.argtypes = [P, I, F, F, P]
            L.orc_texture_sample.restype = None
            L.orc_scene_destroy.argtypes = [P]
            L.orc_scene_destroy.restype = None
            L.orc_trace32.argtypes = [P, P, U32, P, I]
            L.orc_trace32.restype = None
            L.orc_trace32_brute.argtypes = [P, P, U32, P]
            L.orc_trace32_brute.restype = None
            L.orc_trace64.argtypes = [P, P, U32, P, P]
            L.orc_trace64.restype = None
            L.orc_initial_sample.argtypes = [P, P, U32, U32, P, U32, P, P, P, P, P]
            L.orc_initial_sample.restype = U64
            L.orc_initial_sample_cache.argtypes = [P, P, U32, U32, P, U32, P, P, P, P, P, U32, F, P]
            L.orc_initial_sample_cache.restype = U64
            L.orc_mlp_param_count.restype = I
            L.orc_mlp_layer_offset.argtypes = [I, I]
            L.orc_mlp_layer_offset.restype = I
            L.orc_bf16_round.argtypes = [P, U32, P]
            L.orc_bf16_round.restype = None
            L.orc_mlp_encode.argtypes = [P, P, P, P, P, P, P]
            L.orc_mlp_encode.restype = None
            L.orc_mlp_forward.argtypes = [P, P, U32, I, P]
            L.orc_mlp_forward.restype = None
            L.orc_mlp_loss_grad.argtypes = [P, P, P, U32, I, P]
            L.orc_mlp_loss_grad.restype = C.c_double
            L.orc_adam_step.argtypes = [P, P, P, P, U32, C.c_double, C.c_double, C.c_double, C.c_double, U32]
            L.orc_adam_step.restype = None
            L.orc_e4m3_round.argtypes = [P, U32, P, P]
            L.orc_e4m3_round.restype = None
            L.orc_mlp_quantize.argtypes = [P, P, P]
            L.orc_mlp_quantize.restype = None
            L.orc_temporal.argtypes = [U32, U32, U32, U32, P, U32, P, P, P, P, P, P]
            L.orc_temporal.restype = None
            L.orc_spatial.argtypes = [P, U32, U32, U32, U32, U32, F, U32, I, P, U32, P, P, P, P]
            L.orc_spatial.restype = U64
            L.orc_plain.argtypes = [P, P, U32, U32, U32, P, U32, P, P]
            L.orc_plain.restype = U64
            cls._lib = L
        return cls._lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleScene:
    """Oracle copy of a scene (its own FP64-box BVH) plus the pass restatements."""

    def __init__(self, vertices, indices, materials, offsets, textures=()):
        """textures: sequence of (width, height, RGBA8 uint8 array of width*height*4)."""
        self.L = Oracle.lib()
        self.v = np.ascontiguousarray(vertices)
        self.i = np.ascontiguousarray(indices, np.uint32)
        self.m = np.ascontiguousarray(materials)
        self.o = np.ascontiguousarray(offsets, np.uint32).reshape(-1)
        self.h = self.L.orc_scene_create(self.v.ctypes.data, len(self.v), self.i.ctypes.data, len(self.i),
                                         self.m.ctypes.data, len(self.m), self.o.ctypes.data, len(self.o) // 2)
        if not self.h:
            raise ValueError("oracle rejected the scene: an index is out of range for its instance")
        if len(textures):
            info, chunks, first = [], [], 0
            for w, h, rgba in textures:
                info += [first, w, h]
                chunks.append(np.ascontiguousarray(rgba, np.uint8).reshape(-1))
                first += w * h
            self.tex_info = np.array(info, np.uint32)
            self.texels = np.concatenate(chunks)
            if self.L.orc_scene_set_textures(self.h, self.tex_info.ctypes.data, len(textures),
                                             self.texels.ctypes.data, self.texels.nbytes):
                raise ValueError("oracle rejected the textures")

    def texture_sample(self, t: int, u: float, v: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.L.orc_texture_sample(self.h, t, u, v, out.ctypes.data)
        return out

    def close(self):
        if self.h:
            self.L.orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace32(self, rays: np.ndarray, any_hit: bool = False, brute: bool = False) -> np.ndarray:
        from xvk_ray_amd.native import HIT_DTYPE, RAY_DTYPE
        r = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(len(r), HIT_DTYPE)
        if brute:
            self.L.orc_trace32_brute(self.h, r.ctypes.data, len(r), out.ctypes.data)
        else:
            self.L.orc_trace32(self.h, r.ctypes.data, len(r), out.ctypes.data, int(any_hit))
        return out

    def trace64(self, rays_f64: np.ndarray):
        """rays_f64: (n, 8) float64 (origin, tmin, dir, tmax). Returns (ids, t)."""
        r = np.ascontiguousarray(rays_f64, np.float64)
        n = len(r)
        ids = np.zeros(n, np.uint32)
        ts = np.zeros(n, np.float64)
        self.L.orc_trace64(self.h, r.ctypes.data, n, ids.ctypes.data, ts.ctypes.data)
        return ids, ts

    def initial_sample(self, ubo_bytes: bytes, W: int, H: int, pixels=None, cache=None):
        """cache: None, or (training samples per frame, spread_c) of the radiance-cache
        path termination; the result then also holds `train`, the per-pixel
        training choice (uint8)."""
        from xvk_ray_amd.native import SAMPLE_DTYPE
        n = W * H
        samples = np.zeros(n, SAMPLE_DTYPE)
        gpos = np.zeros((n, 4), np.float32)
        gnrm = np.zeros((n, 4), np.float32)
        gmatuv = np.zeros((n, 4), np.float32)
        motion = np.zeros(n, np.uint32)
        ub = np.frombuffer(ubo_bytes, np.uint8).copy()
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        target, spread = cache if cache is not None else (0, 0.0)
        train = np.zeros(n, np.uint8)
        rays = self.L.orc_initial_sample_cache(self.h, ub.ctypes.data, W, H, _ptr(px), 0 if px is None else len(px),
                                               samples.ctypes.data, gpos.ctypes.data, gnrm.ctypes.data,
                                               gmatuv.ctypes.data, motion.ctypes.data, target, spread,
                                               train.ctypes.data)
        return dict(samples=samples, gpos=gpos, gnrm=gnrm, gmatuv=gmatuv, motion=motion, rays=rays, train=train)

    def spatial(self, W, H, frame, reuse, num_samples, radius, m_max, visibility, samples, temporal, spatial, output,
                pixels=None) -> int:
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        return self.L.orc_spatial(self.h, W, H, frame, reuse, num_samples, radius, m_max, visibility, _ptr(px),
                                  0 if px is None else len(px), samples.ctypes.data, temporal.ctypes.data,
                                  spatial.ctypes.data, output.ctypes.data)

    def plain(self, ubo_bytes: bytes, W, H, max_depth, accum, output, pixels=None) -> int:
        ub = np.frombuffer(ubo_bytes, np.uint8).copy()
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        return self.L.orc_plain(self.h, ub.ctypes.data, W, H, max_depth, _ptr(px), 0 if px is None else len(px),
                                accum.ctypes.data, output.ctypes.data)


def temporal(W, H, frame, m_max, samples, motion, gpos_prev, temporal_buf, spatial_buf, output, pixels=None):
    """Temporal pass restatement over whole reference-layout buffers (in place)."""
    L = Oracle.lib()
    px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
    L.orc_temporal(W, H, frame, m_max, _ptr(px), 0 if px is None else len(px), samples.ctypes.data,
                   motion.ctypes.data, gpos_prev.ctypes.data, temporal_buf.ctypes.data, spatial_buf.ctypes.data,
                   output.ctypes.data)


# ---- MLP radiance cache ----
MLP_INPUTS = 96
MLP_PARAMS = 23043


def bf16_round(x) -> np.ndarray:
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(a)
    Oracle.lib().orc_bf16_round(a.ctypes.data, a.size, out.ctypes.data)
    return out


def e4m3_round(x):
    """(codes uint8, values float32) of FP8 E4M3 round-to-nearest-even with saturation."""
    a = np.ascontiguousarray(x, np.float32)
    codes = np.zeros(a.size, np.uint8)
    vals = np.zeros(a.size, np.float32)
    Oracle.lib().orc_e4m3_round(a.ctypes.data, a.size, codes.ctypes.data, vals.ctypes.data)
    return codes, vals


def mlp_quantize(params):
    """(E4M3 codes of the hidden layers' W^T, 22,528 bytes; 5 per-layer scales)."""
    w = np.ascontiguousarray(params, np.float32)
    codes = np.zeros(22528, np.uint8)
    scales = np.zeros(5, np.float32)
    Oracle.lib().orc_mlp_quantize(w.ctypes.data, codes.ctypes.data, scales.ctypes.data)
    return codes, scales


def mlp_layer_offsets(layer: int) -> tuple[int, int]:
    """(weight offset, bias offset) of a layer in the parameter vector."""
    L = Oracle.lib()
    return L.orc_mlp_layer_offset(layer, 0), L.orc_mlp_layer_offset(layer, 1)


def mlp_encode(p, w, n, a, lo, inv_ext) -> np.ndarray:
    args = [np.ascontiguousarray(v, np.float32) for v in (p, w, n, a, lo, inv_ext)]
    out = np.zeros(MLP_INPUTS, np.float32)
    Oracle.lib().orc_mlp_encode(*[v.ctypes.data for v in args], out.ctypes.data)
    return out


def mlp_forward(params, x, mode: int = 1) -> np.ndarray:
    """mode 0: exact FP64 network; mode 1: the kernels' BF16 rounding points;
    mode 2: mode 1 with FP8 E4M3 hidden-layer weights."""
    w = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
    out = np.zeros((len(xx), 3), np.float64)
    Oracle.lib().orc_mlp_forward(w.ctypes.data, xx.ctypes.data, len(xx), mode, out.ctypes.data)
    return out


def mlp_loss_grad(params, x, y, mode: int = 1, grad: bool = True):
    """(loss, gradient or None) of the relative-L2 loss."""
    w = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
    yy = np.ascontiguousarray(y, np.float32).reshape(-1, 3)
    g = np.zeros(MLP_PARAMS, np.float64) if grad else None
    loss = Oracle.lib().orc_mlp_loss_grad(w.ctypes.data, xx.ctypes.data, yy.ctypes.data, len(xx), mode, _ptr(g))
    return float(loss), g


def adam_step(w, m, v, g, lr, beta1, beta2, eps, t) -> None:
    """In place on float64 arrays."""
    for a in (w, m, v, g):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    Oracle.lib().orc_adam_step(w.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, len(w), lr, beta1, beta2,
                               eps, t)
