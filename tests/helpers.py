"""Shared set-up for the tests: scenes, cameras, ray batches, comparisons."""
from __future__ import annotations

import sys

import numpy as np

from xvk_ray_amd import native as N

CORNELL_EYE = (0.5, 0.5, -1.2)
CORNELL_TARGET = (0.5, 0.5, 0.5)


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def cornell_scene(seed: int = 1) -> N.Scene:
    return N.Scene.cornell(seed)


def small_terrain(seed: int = 1, grid: int = 128, spheres: int = 32, lights: int = 8) -> N.Scene:
    return N.Scene.terrain(seed, grid, spheres, lights)


def look_view(eye, target) -> np.ndarray:
    cam = N.Camera.static(eye, target)
    v = cam.view(0)
    cam.close()
    return v


def random_rays(n: int, lo, hi, seed: int, tmin: float = 1e-4, tmax: float = 1e4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    r = np.zeros(n, N.RAY_DTYPE)
    r["origin"] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["direction"] = d.astype(np.float32)
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def rays_f64(r: np.ndarray) -> np.ndarray:
    out = np.zeros((len(r), 8), np.float64)
    out[:, 0:3] = r["origin"]
    out[:, 3] = r["tmin"]
    out[:, 4:7] = r["direction"]
    out[:, 7] = r["tmax"]
    return out


def oracle_scene(scene: N.Scene):
    from oracle.binding import OracleScene
    v, i, m, o = scene.arrays()
    return OracleScene(v, i, m, o, scene.textures())


def bytes_equal_rows(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-entry bitwise equality of two structured / plain arrays of the same dtype."""
    ab = np.ascontiguousarray(a).view(np.uint8).reshape(len(a), -1)
    bb = np.ascontiguousarray(b).view(np.uint8).reshape(len(b), -1)
    return np.all(ab == bb, axis=1)


def describe_mismatch(name: str, a: np.ndarray, b: np.ndarray, eq: np.ndarray, k: int = 3) -> str:
    bad = np.nonzero(~eq)[0]
    lines = [f"{name}: {len(bad)} of {len(eq)} entries differ"]
    for j in bad[:k]:
        lines.append(f"  [{j}] gpu={a[j]!r}\n       ref={b[j]!r}")
    return "\n".join(lines)
