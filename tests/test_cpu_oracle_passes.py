"""CPU tests of the oracle's pass restatements (no GPU).

Pins each pass by closed forms, special cases and invariants:
  * TraceRay: the oracle BVH equals brute force; FP32 equals the FP64 reference
    on >= 99.99% of rays; rays from inside a closed icosphere never escape
    (watertightness across shared edges and vertices).
  * initial sample: hand-computable two-surface scene (floor under a light).
  * temporal: no history -> M = 1; M above the cap -> w_sum rescaled;
    invalid pixels -> initial reservoir in both slots and no output write.
  * spatial: neighbours identical to the centre -> Jacobian 1 and M summed;
    black / white early outs.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import (CORNELL_EYE, CORNELL_TARGET, bytes_equal_rows, cornell_scene, look_view, oracle_scene,
                           random_rays, rays_f64)
from xvk_ray_amd import native as N


def test_oracle_bvh_equals_brute_force():
    orc = oracle_scene(cornell_scene())
    rays = random_rays(4000, 0.02, 0.98, seed=3)
    assert bytes_equal_rows(orc.trace32(rays), orc.trace32(rays, brute=True)).all()


def test_fp32_trace_matches_fp64_reference():
    scene = N.Scene.terrain(4, 128, 16, 8)
    orc = oracle_scene(scene)
    rays = random_rays(20000, 0.0, 0.5, seed=4)
    h32 = orc.trace32(rays)
    ids, ts = orc.trace64(rays_f64(rays))
    same = h32["triangle"] == ids
    assert same.mean() >= 0.9999
    hit = ids != 0xFFFFFFFF
    assert np.allclose(h32["t"][hit & same], ts[hit & same], rtol=1e-5)


def test_watertight_inside_closed_sphere():
    v, i, m, o = cornell_scene().arrays()
    # instance 2 of the Cornell scene is the icosphere (closed mesh)
    sphere = N.Scene.from_arrays(v[o[5]:], i[o[4]:] , m, np.array([0, 0], np.uint32))
    orc = oracle_scene(sphere)
    rng = np.random.default_rng(9)
    n = 30000
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = (0.5 + rng.uniform(-0.05, 0.05, (n, 3))).astype(np.float32)
    # aim at sphere vertices and edge midpoints: the hardest places to leak
    vp = v[o[5]:]["pos"].astype(np.float64)
    tgt = vp[rng.integers(0, len(vp), n)]
    d = tgt - rays["origin"]
    rays["direction"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"], rays["tmax"] = 0.0, 1e4
    h = orc.trace32(rays)
    assert (h["t"] > 0).all(), f"{(h['t'] <= 0).sum()} rays leaked"


def _floor_light_scene():
    """A floor quad (y=0, albedo 0.5) under an emissive quad (y=1, radiance 8)."""
    verts = np.zeros(8, N.VERTEX_DTYPE)
    L = 1000.0  # wide enough that a bounce off the floor misses the light with probability ~1e-6
    verts["pos"] = [[-L, 0, -L], [L, 0, -L], [L, 0, L], [-L, 0, L],
                    [-L, 1, -L], [L, 1, -L], [L, 1, L], [-L, 1, L]]
    verts["normal"][:4] = [0, 1, 0]
    verts["normal"][4:] = [0, -1, 0]
    verts["material_index"][:4] = 0
    verts["material_index"][4:] = 1
    mats = np.zeros(2, N.MATERIAL_DTYPE)
    mats["base_color"] = [[0.5, 0.5, 0.5, 1], [2, 2, 2, 1]]
    mats["base_color_texture_index"] = -1
    mats["material_model"] = [N.MATERIAL_LAMBERTIAN, N.MATERIAL_DIFFUSE_LIGHT]
    idx = np.array([0, 1, 2, 0, 2, 3, 0, 1, 2, 0, 2, 3], np.uint32)  # instance-local, as AScene stores them
    return N.Scene.from_arrays(verts, idx, mats, np.array([0, 0, 6, 4], np.uint32))


def test_initial_sample_floor_under_light():
    scene = _floor_light_scene()
    orc = oracle_scene(scene)
    W, H = 16, 12
    view = look_view((0, 0.5, 0), (0.0, 0.0, -1.0))  # looking down at the floor
    ubo = N.make_ubo(view, view, 60.0, W, H, bounces=4, frame=1)
    out = orc.initial_sample(N.ubo_to_bytes(ubo), W, H)
    s = out["samples"]
    floor = s["material_id"] == 0
    assert floor.any()
    # every floor pixel: bounce 1 goes up and must hit the light (infinite planes)
    # (interpolated positions: 1*b0 + 1*b1 + 1*b2 rounds to within an ulp of 1)
    assert np.allclose(s["x_sample"][floor][:, 1], 1.0, rtol=0, atol=4e-7)
    assert np.allclose(s["n_sample"][floor], [0, -1, 0])
    assert np.all(s["l_out"][floor] == np.float32(8.0))  # baseColor * 4, throughput 1 at bounce 1
    assert np.all(s["x_view"][floor][:, 1] == 0.0)  # 0*b0 + 0*b1 + 0*b2 is exact
    assert np.allclose(s["f"][floor], 0.5 / np.pi, rtol=1e-6)
    # p_q is the cosine pdf of the chosen direction: cos/pi, the bounce ray
    # leaving x_v + 0.01 * geometric normal (InitialSample.rgen:175-178)
    d = s["x_sample"][floor] - (s["x_view"][floor] + np.float32([0, 0.01, 0]))
    cos = d[:, 1] / np.linalg.norm(d, axis=1)
    # (positions up to ~1000 units from the origin carry ~1e-5 relative FP32 error)
    assert np.allclose(s["p_q"][floor], cos / np.pi, rtol=2e-3)
    assert np.all(out["gpos"][floor][:, 3] == 1.0)
    assert out["rays"] == 2 * floor.sum() + (~floor).sum()


def _synthetic_frame(W, H, seed, invalid_frac=0.05):
    """Kernel-isolation inputs: positions U[-1,1]^3, normals on S^2, L_out U[0,5]^3,
    p_q U[0.01,1/pi], M U{0..20}, w_sum U[0,10], motion N(0, 2 px)."""
    rng = np.random.default_rng(seed)
    n = W * H

    def unit(k):
        v = rng.normal(size=(k, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    s = np.zeros(n, N.SAMPLE_DTYPE)
    s["x_view"] = rng.uniform(-1, 1, (n, 3))
    s["n_view"] = unit(n)
    s["x_sample"] = rng.uniform(-1, 1, (n, 3))
    s["n_sample"] = unit(n)
    s["l_out"] = rng.uniform(0, 5, (n, 3))
    s["f"] = rng.uniform(0, 0.3, (n, 3))
    s["p_q"] = rng.uniform(0.01, 1 / np.pi, n)
    s["material_id"] = rng.integers(0, 3, n)
    s["rand_seed"] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    bad = rng.random(n) < invalid_frac
    s["n_sample"][bad] = 0
    s["l_out"][bad] = 0
    t = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    t["s"] = np.concatenate([s, s])
    t["m"] = rng.integers(0, 21, 2 * n)
    t["w_sum"] = rng.uniform(0, 10, 2 * n)
    t["w"] = rng.uniform(0, 2, 2 * n)
    gpos = np.zeros((n, 4), np.float32)
    gpos[:, :3] = s["x_view"]
    gpos[:, 3] = (~bad).astype(np.float32)
    mv = rng.normal(0, 2, (n, 2)) / np.array([W, H])
    h = mv.astype(np.float16).view(np.uint16).astype(np.uint32)
    motion = (h[:, 0] | (h[:, 1] << 16)).astype(np.uint32)
    return s, t, gpos, motion, bad


def test_temporal_special_cases():
    from oracle.binding import temporal
    W, H = 40, 30
    n = W * H
    s, t, gpos, motion, bad = _synthetic_frame(W, H, 1)
    motion[:] = 0  # identity reprojection
    spatial = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    out = np.full(n, 0x12345678, np.uint32)
    frame = 3  # ping = 1: current slot n.., previous slot 0..n-1
    gpos[: n // 2, 3] = 0  # first half: no history
    t0 = t.copy()
    temporal(W, H, frame, 20, s, motion, gpos, t, spatial, out)
    cur = t[n:]
    hist = gpos[:, 3] != 0
    # invalid pixels: initial reservoir in both current slots, output untouched
    assert np.all(cur["m"][bad] == 0) and np.all(cur["w_sum"][bad] == 0)
    assert np.all(cur["s"]["material_id"][bad] == 0xFFFFFFFF)
    assert np.all(spatial[n:]["m"][bad] == 0)
    assert np.all(out[bad] == 0x12345678)
    # no history (valid pixel, no history at its own position): M = 1
    nh = (~bad) & (~hist)
    assert np.all(cur["m"][nh] == 1)
    assert np.all(out[nh] != 0x12345678)
    # history: M = min(1 + M_prev, 20); above the cap w_sum is rescaled by 20/M
    wh = (~bad) & hist
    m_prev = t0[:n]["m"][wh]
    assert np.array_equal(cur["m"][wh], np.minimum(1 + m_prev, 20))
    capped = (1 + m_prev) > 20
    assert capped.any()
    # previous slot untouched
    assert bytes_equal_rows(t[:n], t0[:n]).all()


def test_spatial_identical_neighbours():
    """A constant image: every neighbour equals the centre, so each similar
    neighbour merges with Jacobian 1 and the M sum is 1 + K*m (capped)."""
    W, H = 24, 16
    n = W * H
    scene = cornell_scene()
    orc = oracle_scene(scene)
    s = np.zeros(n, N.SAMPLE_DTYPE)
    s["x_view"] = [0.1, 0.2, 0.3]
    s["n_view"] = [0, 1, 0]
    s["x_sample"] = [0.1, 1.2, 0.3]
    s["n_sample"] = [0, -1, 0]
    s["l_out"] = [1, 2, 2]
    s["f"] = 0.1
    s["material_id"] = 1
    t = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    t["s"] = np.concatenate([s, s])
    t["m"] = 2
    t["w"] = 0.5
    t["w_sum"] = 3.0
    sp = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    out = np.zeros(n, np.uint32)
    frame = 2  # read slot 0, write slot 1
    orc.spatial(W, H, frame, 1, 6, 5.0, 300, 0, s, t, sp, out)
    w = sp[n:]
    assert np.all(w["m"] == 2 + 6 * 2)
    # w_sum = sum of p_hat * W * m over centre + 6 neighbours, p_hat = |L| = 3
    assert np.allclose(w["w_sum"], 7 * (3.0 * 0.5 * 2), rtol=1e-6)
    assert np.allclose(w["w"], w["w_sum"] / (w["m"] * 3.0), rtol=1e-6)
    # early outs
    s2 = s.copy()
    s2["n_view"][:5] = 0
    s2["n_sample"][5:9] = 0
    out2 = np.zeros(n, np.uint32)
    sp2 = np.zeros(2 * n, N.RESERVOIR_DTYPE)
    orc.spatial(W, H, frame, 1, 6, 5.0, 300, 0, s2, t, sp2, out2)
    assert np.all(out2[:5] == 0xFF000000) and np.all(out2[5:9] == 0xFFFFFFFF)
    assert np.all(sp2[n:n + 9]["m"] == 0)  # no reservoir write on early out


def test_plain_pt_floor_under_light_closed_form():
    """Plain PT, max depth 2: a floor pixel's path is floor -> light, so its
    radiance is (albedo/pi * cos / (cos/pi)) * 8 = 0.5 * 8 = 4 exactly up to rounding."""
    scene = _floor_light_scene()
    orc = oracle_scene(scene)
    W, H = 8, 6
    view = look_view((0, 0.5, 0), (0.0, 0.0, -1.0))
    ubo = N.make_ubo(view, view, 60.0, W, H, spp=1, total=1, bounces=10, frame=1)
    accum = np.zeros((W * H, 4), np.float32)
    out = np.zeros(W * H, np.uint32)
    orc.plain(N.ubo_to_bytes(ubo), W, H, 2, accum, out)
    # pixels whose camera ray reaches the light directly (top rows look above the horizon) see 8
    floor = ~np.isclose(accum[:, 0], 8.0)
    assert floor.sum() > 0.8 * W * H
    assert np.allclose(accum[floor, :3], 4.0, rtol=1e-5)
