"""CPU tests pinning the oracle's scalar building blocks (no GPU).

The reference ships no tests, fixtures or known-answer vectors, so each
oracle function is pinned by what the mathematics fixes:
  * RNG: the PCG state recurrence is an LCG with a closed-form jump-ahead;
    the RXS-M-XS output step is a bijection (inverted here step by step);
    the seed hash is checked against committed golden vectors
    (tests/golden/rng_kat.json, produced by tests/golden/make_rng_kat.py, a
    separate big-integer evaluation of random.hlsl:1-10).
  * cos/sin: against FP64 numpy within 2 ulp.
  * FP16 store: against numpy's IEEE round-to-nearest-even conversion.
  * ONB, cosine hemisphere sampling, Lambertian f and pdf: orthonormality,
    chi-square of the cosine distribution, f = albedo/pi, pdf = cos/pi.
  * ACES / output conversion: closed-form values.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
M32 = 0xFFFFFFFF


def test_rng_seed_hash_golden(oracle_lib):
    kat = json.loads((GOLDEN / "rng_kat.json").read_text())
    for case in kat["seeds"]:
        got = oracle_lib.orc_init_seed(case["x"], case["y"], case["w"], case["frame"])
        assert got == case["seed"], case


def test_rng_stream_golden(oracle_lib):
    kat = json.loads((GOLDEN / "rng_kat.json").read_text())
    for case in kat["streams"]:
        n = len(case["words"])
        words = np.zeros(n, np.uint32)
        floats = np.zeros(n, np.float32)
        oracle_lib.orc_random_stream(case["state"], n, words.ctypes.data, floats.ctypes.data)
        assert words.tolist() == case["words"]
        assert np.array_equal(floats, np.array(case["floats"], np.float32))


def _lcg_jump(state: int, n: int) -> int:
    """state_n of s' = a*s + c (mod 2^32) by binary exponentiation of the affine map."""
    a, c = 747796405, 2891336453
    acc_a, acc_c = 1, 0
    cur_a, cur_c = a, c
    while n:
        if n & 1:
            acc_a, acc_c = (cur_a * acc_a) & M32, (cur_a * acc_c + cur_c) & M32
        cur_a, cur_c = (cur_a * cur_a) & M32, (cur_a * cur_c + cur_c) & M32
        n >>= 1
    return (acc_a * state + acc_c) & M32


def _invert_output(word: int) -> int:
    """Inverse of RXS-M-XS: recovers the LCG state from one output word."""
    # undo w ^= w >> 22
    w = word ^ (word >> 22)
    # undo multiplication by 277803737 (odd => invertible mod 2^32)
    w = (w * pow(277803737, -1, 1 << 32)) & M32
    # undo x ^= x >> (x >> 28) + 4 ; the top 4 bits are unchanged by it
    shift = (w >> 28) + 4
    x = w
    for _ in range(32 // shift + 1):
        x = w ^ (x >> shift)
    return x


def test_rng_state_recurrence_and_output_bijection(oracle_lib):
    rng = np.random.default_rng(0)
    for start in rng.integers(0, 2**32, size=20, dtype=np.uint64):
        start = int(start)
        n = 64
        words = np.zeros(n, np.uint32)
        floats = np.zeros(n, np.float32)
        oracle_lib.orc_random_stream(start, n, words.ctypes.data, floats.ctypes.data)
        for k in range(n):
            state_k = _lcg_jump(start, k + 1)
            assert _invert_output(int(words[k])) == state_k
            assert floats[k] == np.float32((int(words[k]) & 0xFFFFFF) / 16777216.0)


def test_rng_floats_uniform(oracle_lib):
    n = 200000
    words = np.zeros(n, np.uint32)
    floats = np.zeros(n, np.float32)
    oracle_lib.orc_random_stream(12345, n, words.ctypes.data, floats.ctypes.data)
    assert floats.min() >= 0.0 and floats.max() < 1.0
    hist, _ = np.histogram(floats, bins=64, range=(0, 1))
    exp = n / 64
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5


def test_sincos_accuracy(oracle_lib):
    x = np.linspace(0, 2 * np.pi, 200001, dtype=np.float32)
    x = np.concatenate([x, (np.arange(1 << 16, dtype=np.float32) / 65536.0) * np.float32(6.28318530718)])
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    oracle_lib.orc_sincos(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    xs = x.astype(np.float64)
    for got, ref in ((s, np.sin(xs)), (c, np.cos(xs))):
        ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - ref)
        # relative ulp bound, plus absolute slack near zeros of sin/cos
        assert np.all(err <= 2 * ulp + 2e-8), float((err / ulp).max())


def test_half_store_matches_ieee_rne(oracle_lib):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.normal(scale=1e-3, size=20000), rng.normal(scale=1.0, size=20000),
        rng.uniform(-70000, 70000, size=2000), [0.0, -0.0, 6.1e-5, 5.96e-8, 2.98e-8, 1e-9, 65504.0, 65520.0, 1e6],
    ]).astype(np.float32)
    to_half = oracle_lib.orc_to_half
    from_half = oracle_lib.orc_from_half
    for v in vals:
        h = to_half(C.c_float(v))
        ref = np.array([v], np.float32).astype(np.float16).view(np.uint16)[0]
        assert h == ref, (v, hex(h), hex(ref))
        back = from_half(h)
        assert np.float32(back) == np.float32(np.array([h], np.uint16).view(np.float16)[0].astype(np.float32))


def test_onb_orthonormal(oracle_lib):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        b1 = np.zeros(3, np.float32)
        b2 = np.zeros(3, np.float32)
        oracle_lib.orc_onb(n.ctypes.data, b1.ctypes.data, b2.ctypes.data)
        M = np.stack([b1, b2, n]).astype(np.float64)
        assert np.allclose(M @ M.T, np.eye(3), atol=2e-6)
        assert np.allclose(np.cross(b1, b2), n, atol=2e-6)  # right-handed frame


def test_lambert_sampling_cosine_distribution(oracle_lib):
    n = np.array([0.3, -0.5, 0.8], np.float64)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    albedo = np.array([0.2, 0.5, 0.9], np.float32)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    cs = C.c_float()
    seed_out = C.c_uint32()
    N = 40000
    cos_vals = np.zeros(N)
    for k in range(N):
        oracle_lib.orc_sample_lambert(n.ctypes.data, albedo.ctypes.data, k * 7919 + 1, wi.ctypes.data,
                                      f.ctypes.data, C.byref(pdf), C.byref(cs), C.byref(seed_out))
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        c = float(np.dot(wi, n))
        assert abs(c - cs.value) < 1e-5
        assert pdf.value == np.float32(np.float32(cs.value) * np.float32(1 / 3.14159265359))
        assert np.array_equal(f, albedo * np.float32(1 / 3.14159265359))
        cos_vals[k] = c
    # cos(theta) of a cosine-weighted hemisphere: P(cos^2 <= u) = u
    u = cos_vals ** 2
    hist, _ = np.histogram(u, bins=40, range=(0, 1))
    exp = N / 40
    assert ((hist - exp) ** 2 / exp).sum() < 90  # 39 dof
    assert cos_vals.min() > 0


def test_aces_and_output_closed_forms(oracle_lib):
    aces = oracle_lib.orc_aces
    assert aces(0.0) == 0.0
    assert abs(aces(1.0) - 2.54 / 3.16) < 1e-6
    assert aces(1e6) == 1.0
    res = oracle_lib.orc_resolve
    black = np.zeros(3, np.float32)
    assert res(black.ctypes.data, 1.0) == 0xFF000000
    big = np.full(3, 1e6, np.float32)
    assert res(big.ctypes.data, 0.0) == 0x00FFFFFF
    # pure red channel lands in byte 2 (B8G8R8A8)
    red = np.array([1e6, 0, 0], np.float32)
    assert res(red.ctypes.data, 1.0) == 0xFFFF0000
    x = np.array([0.5, 0.5, 0.5], np.float32)
    tone = (0.5 * (2.51 * 0.5 + 0.03)) / (0.5 * (2.43 * 0.5 + 0.59) + 0.14)
    expect = round(tone ** (1 / 2.2) * 255)  # 205
    got = res(x.ctypes.data, 1.0) & 0xFF
    assert abs(got - expect) <= 1


@pytest.mark.parametrize("w_new", [0.0, 0.25, 3.0])
def test_update_reservoir_closed_form(oracle_lib, w_new):
    from xvk_ray_amd.native import RESERVOIR_DTYPE, SAMPLE_DTYPE
    r = np.zeros(1, RESERVOIR_DTYPE)
    r["w_sum"] = 1.5
    r["m"] = 4
    r["s"]["material_id"] = 7
    s = np.zeros(1, SAMPLE_DTYPE)
    s["material_id"] = 9
    seed = C.c_uint32(1234)
    xi_words = np.zeros(1, np.uint32)
    xi = np.zeros(1, np.float32)
    oracle_lib.orc_random_stream(1234, 1, xi_words.ctypes.data, xi.ctypes.data)
    oracle_lib.orc_update_reservoir(r.ctypes.data, s.ctypes.data, w_new, C.byref(seed))
    assert r["w_sum"][0] == np.float32(1.5 + w_new)
    assert r["m"][0] == 5
    take = xi[0] < np.float32(w_new) / max(np.float32(1e-10), np.float32(1.5 + w_new))
    assert r["s"]["material_id"][0] == (9 if take else 7)
    assert seed.value == _lcg_jump(1234, 1)  # exactly one draw, taken or not


def test_merge_reservoir_m_and_weight(oracle_lib):
    from xvk_ray_amd.native import RESERVOIR_DTYPE
    dst = np.zeros(1, RESERVOIR_DTYPE)
    dst["m"] = 3
    dst["w_sum"] = 2.0
    src = np.zeros(1, RESERVOIR_DTYPE)
    src["m"] = 17
    src["w"] = 0.5
    seed = C.c_uint32(99)
    oracle_lib.orc_merge_reservoir(dst.ctypes.data, src.ctypes.data, 2.0, C.byref(seed))
    assert dst["m"][0] == 20  # M0 + src.m, not M0 + 1
    assert dst["w_sum"][0] == np.float32(2.0 + 2.0 * 0.5 * 17)


def test_streaming_ris_selects_proportionally(oracle_lib):
    """Weighted reservoir sampling keeps candidate i with probability w_i / sum(w)."""
    from xvk_ray_amd.native import RESERVOIR_DTYPE, SAMPLE_DTYPE
    w = np.array([0.5, 2.0, 0.0, 1.0, 4.5], np.float32)
    cands = np.zeros(len(w), SAMPLE_DTYPE)
    cands["material_id"] = np.arange(len(w))
    trials = 20000
    counts = np.zeros(len(w))
    for t in range(trials):
        r = np.zeros(1, RESERVOIR_DTYPE)
        r["s"]["material_id"] = 0xFFFFFFFF
        seed = C.c_uint32(t * 2654435761 & M32)
        for k in range(len(w)):
            oracle_lib.orc_update_reservoir(r.ctypes.data, cands[k:k + 1].ctypes.data, float(w[k]), C.byref(seed))
        counts[r["s"]["material_id"][0]] += 1
    p = w / w.sum()
    assert counts[2] == 0
    nz = p > 0
    chi2 = ((counts[nz] - trials * p[nz]) ** 2 / (trials * p[nz])).sum()
    assert chi2 < 20  # 3 dof
