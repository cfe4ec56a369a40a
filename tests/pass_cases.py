"""Hand-built scenes and pass inputs whose results the mathematics fixes.

Shared by the CPU closed-form tests of the oracle and the GPU tests that run
the same inputs through the HIP passes.
"""
from __future__ import annotations

import numpy as np

from xvk_ray_amd import native as N


def quads_scene(quads, materials, textures=()) -> N.Scene:
    """quads: list of (4 corner positions, normal, material index, uv of the corners or None)."""
    verts = np.zeros(4 * len(quads), N.VERTEX_DTYPE)
    idx = []
    for k, (corners, normal, mat, uv) in enumerate(quads):
        verts["pos"][4 * k:4 * k + 4] = corners
        verts["normal"][4 * k:4 * k + 4] = normal
        verts["material_index"][4 * k:4 * k + 4] = mat
        if uv is not None:
            verts["tex_coord"][4 * k:4 * k + 4] = uv
        idx += [4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 0, 4 * k + 2, 4 * k + 3]
    return N.Scene.from_arrays(verts, np.array(idx, np.uint32), materials, np.array([0, 0], np.uint32), textures)


def materials(colors, models, textures=None) -> np.ndarray:
    m = np.zeros(len(colors), N.MATERIAL_DTYPE)
    m["base_color"] = [list(c) + [1.0] for c in colors]
    m["base_color_texture_index"] = -1 if textures is None else textures
    m["material_model"] = models
    return m


def floor_light_textured() -> N.Scene:
    """Floor y=0 (Lambertian, base colour 0.5, texture 1 = white) under a light
    y=1 (base colour 2 -> radiance 8, texture 0 = black)."""
    L = 1000.0
    floor = ([[-L, 0, -L], [L, 0, -L], [L, 0, L], [-L, 0, L]], [0, 1, 0], 0, [[0, 0], [1, 0], [1, 1], [0, 1]])
    light = ([[-L, 1, -L], [L, 1, -L], [L, 1, L], [-L, 1, L]], [0, -1, 0], 1, [[0, 0], [1, 0], [1, 1], [0, 1]])
    mats = materials([(0.5, 0.5, 0.5), (2.0, 2.0, 2.0)], [N.MATERIAL_LAMBERTIAN, N.MATERIAL_DIFFUSE_LIGHT],
                     textures=[1, 0])
    tex = [(1, 1, np.array([0, 0, 0, 255], np.uint8)), (1, 1, np.array([255, 255, 255, 255], np.uint8))]
    return quads_scene([floor, light], mats, tex)


def wall_scene() -> N.Scene:
    """One Lambertian wall at z = -1 spanning x, y in [-10, 10]."""
    w = ([[-10, -10, -1], [10, -10, -1], [10, 10, -1], [-10, 10, -1]], [0, 0, 1], 0, None)
    return quads_scene([w], materials([(0.5, 0.5, 0.5)], [N.MATERIAL_LAMBERTIAN]))


def translation_view(eye) -> np.ndarray:
    """Column-major view matrix of a camera at `eye` looking down -z with +y up:
    the identity rotation followed by a translation by -eye."""
    v = np.eye(4, dtype=np.float32)
    v[3, :3] = -np.asarray(eye, np.float32)  # column 3 of the column-major matrix
    return v.reshape(-1)


def occluder_scene() -> N.Scene:
    """A 0.08 x 0.08 quad at y = 0.5 centred on (-0.05, 0.5, 0): it blocks the
    segment from (-0.1, 0.01, 0) to (0, 1, 0) and not the one from (0.1, 0.01, 0)."""
    q = ([[-0.09, 0.5, -0.04], [-0.01, 0.5, -0.04], [-0.01, 0.5, 0.04], [-0.09, 0.5, 0.04]], [0, -1, 0], 0, None)
    return quads_scene([q], materials([(0.5, 0.5, 0.5)], [N.MATERIAL_LAMBERTIAN]))


def spatial_layout(W, H, centre, x_view_centre, x_view_other, res_x_view, x_sample, n_sample, l_out=(1.0, 2.0, 2.0),
                   W_res=0.5, m_res=2):
    """Samples and read-slot temporal reservoirs (frame 2 reads slot 0) for a
    spatial pass: every pixel's visible point is x_view_other except `centre`'s;
    every temporal reservoir holds the same sample (x_view res_x_view, x_sample,
    n_sample, l_out) with weight W_res and count m_res."""
    n = W * H
    s = np.zeros(n, N.SAMPLE_DTYPE)
    s["x_view"] = x_view_other
    s["x_view"][centre] = x_view_centre
    s["n_view"] = [0, 1, 0]
    s["x_sample"] = x_sample
    s["n_sample"] = n_sample
    s["l_out"] = l_out
    s["f"] = 0.1
    s["p_q"] = 0.2
    s["material_id"] = 1
    t = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    rs = s.copy()
    rs["x_view"] = res_x_view
    t["s"] = np.concatenate([rs, rs])
    t["m"] = m_res
    t["w"] = W_res
    t["w_sum"] = 3.0
    return s, t


def jacobian(x_view_orig, x_view_curr, x_sample, n_sample) -> float:
    """(cos_curr * d2_orig) / (cos_orig * d2_curr) in float64, clamped to [0.1, 10]
    (the reconnection Jacobian of Spatial.rgen:113-133)."""
    xs, ns = np.asarray(x_sample, np.float64), np.asarray(n_sample, np.float64)
    vo, vc = np.asarray(x_view_orig, np.float64) - xs, np.asarray(x_view_curr, np.float64) - xs
    d2o, d2c = vo @ vo, vc @ vc
    co, cc = max(0.0, ns @ vo / np.sqrt(d2o)), max(0.0, ns @ vc / np.sqrt(d2c))
    return float(np.clip(cc * d2o / (co * d2c), 0.1, 10.0))


# Three reconnection geometries for the centre pixel (8, 8) of a 16 x 16 image:
# name -> (centre x_view, other pixels' x_view, x_sample, n_sample).
JACOBIAN_CASES = {
    "inside": ((0.2, 0.0, 0.1), (0.0, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, -1.0, 0.0)),   # J = 0.76
    "low_clamp": ((0.25, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.02, 0.0), (0.0, -1.0, 0.0)),  # J = 5e-4 -> 0.1
    "high_clamp": ((0.25, 0.0, 0.0), (0.0, 0.0, 0.0), (0.25, 0.02, 0.0), (0.0, -1.0, 0.0)),  # J = 1973 -> 10
}
