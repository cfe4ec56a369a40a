// MLP radiance cache: parameter layout, input encoding, BF16 helpers.
//
// No counterpart in the reference.  The network maps a path vertex to the
// radiance leaving it: 96 inputs (72 frequency-encoded position values, view
// direction, normal, albedo, roughness, 13 zero pads) -> 5 hidden layers of
// 64 ReLU units -> 3 linear outputs (RGB).  23,043 parameters.  BF16 weights
// and activations on MFMA with FP32 accumulation; FP32 master weights and
// Adam state.
#pragma once
#include "xr_math.h"

namespace xr {

constexpr int kMlpIn = 96;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;
constexpr int kMlpLayers = 6;  // 5 hidden + output
constexpr int kMlpBands = 12;  // frequency bands per position axis
constexpr int kMlpFeatures = 3 * 2 * kMlpBands + 11;  // 83 used inputs

// FP32 parameter vector: for layer l, W_l is [in_l][out_l] row-major
// (W[k][j] = weight from input k to output j) followed by its bias [out_l].
XR_HD int mlp_in(int l) { return l == 0 ? kMlpIn : kMlpHidden; }
XR_HD int mlp_out(int l) { return l == kMlpLayers - 1 ? kMlpOut : kMlpHidden; }
XR_HD int mlp_woff(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += mlp_in(i) * mlp_out(i) + mlp_out(i);
    return o;
}
XR_HD int mlp_boff(int l) { return mlp_woff(l) + mlp_in(l) * mlp_out(l); }
constexpr int kMlpParams = 96 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 3 + 3;
static_assert(kMlpParams == 23043, "parameter count");

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet).
XR_HD uint16_t bf16_bits(float f) {
    uint32_t u = as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
XR_HD float bf16_float(uint16_t h) { return as_float((uint32_t)h << 16); }
XR_HD float bf16_round(float f) { return bf16_float(bf16_bits(f)); }

// FP32 -> FP8 E4M3 (the OCP "E4M3FN" format: 1 sign, 4 exponent bits with
// bias 7, 3 mantissa bits, no infinities, largest finite 448, NaN
// S.1111.111), round to nearest even, saturating at +-448.
XR_HD uint8_t e4m3_bits(float f) {
    const uint32_t u = as_uint(f);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const float a = fabsf(f);
    if (!(a == a)) return (uint8_t)(sign | 0x7Fu);
    if (a < 0.015625f) {  // below 2^-6: multiples of 2^-9 (8 * 2^-9 is the smallest normal, code 0x08)
        const uint32_t q = (uint32_t)rintf(a * 512.0f);
        return (uint8_t)(sign | q);
    }
    const uint32_t ua = as_uint(a);
    int e = (int)((ua >> 23) & 0xFFu) - 127;
    uint32_t keep = (ua & 0x7FFFFFu) >> 20;
    const uint32_t rest = ua & 0xFFFFFu;
    if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) ++keep;
    if (keep == 8u) {
        keep = 0u;
        ++e;
    }
    if (e > 8 || (e == 8 && keep > 6u)) return (uint8_t)(sign | 0x7Eu);  // 448
    return (uint8_t)(sign | ((uint32_t)(e + 7) << 3) | keep);
}
XR_HD float e4m3_float(uint8_t b) {
    const uint32_t E = (b >> 3) & 0xFu, m = b & 7u;
    float v;
    if (E == 0u) v = (float)m * 0.001953125f;  // m * 2^-9
    else if (E == 15u && m == 7u) v = as_float(0x7FC00000u);
    else v = as_float(((E + 120u) << 23) | (m << 20));
    return (b & 0x80u) ? -v : v;
}
// Per-layer power-of-two scale of the FP8 weights: the smallest 2^k with
// max|W| / 2^k <= 448 (1 for an all-zero layer).
XR_HD float e4m3_layer_scale(float maxabs) {
    if (!(maxabs > 0.0f)) return 1.0f;
    int k;
    const float m = frexpf(maxabs / 448.0f, &k);  // maxabs / 448 = m * 2^k, m in [0.5, 1)
    return ldexpf(1.0f, m == 0.5f ? k - 1 : k);
}

// The 24 frequency features of one normalised coordinate u: out[2k] =
// sin(2^k pi u), out[2k+1] = cos(2^k pi u).
XR_HD void mlp_encode_axis(float u, float* out) {
    float scale = XR_PI;
    for (int k = 0; k < kMlpBands; ++k) {
        float s, c;
        sincos_xr(scale * u, &s, &c);
        out[2 * k] = s;
        out[2 * k + 1] = c;
        scale = scale * 2.0f;
    }
}

// Inputs [72, 96): w, n, a, roughness 0, zero padding.
XR_HD void mlp_encode_tail(v3 w, v3 n, v3 a, float* out) {
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
    out[3] = n.x; out[4] = n.y; out[5] = n.z;
    out[6] = a.x; out[7] = a.y; out[8] = a.z;
    out[9] = 0.0f;
    for (int i = 10; i < kMlpIn - 6 * kMlpBands; ++i) out[i] = 0.0f;
}

// The 96 network inputs of a vertex at position p (in world units, scaled to
// [0,1]^3 by the scene box lo / inv_ext), outgoing direction w (unit,
// pointing away from the vertex), shading normal n, albedo a:
//   [0, 72)  sin(2^k pi u), cos(2^k pi u) for k = 0..11, u each of the three
//            normalised coordinates (axis-major: x's 24 values, then y, z)
//   [72, 75) w   [75, 78) n   [78, 81) a   [81] roughness (0: Lambertian)
//   [82, 96) zero
XR_HD void mlp_encode(v3 p, v3 w, v3 n, v3 a, v3 lo, v3 inv_ext, float* out) {
    const float u[3] = {(p.x - lo.x) * inv_ext.x, (p.y - lo.y) * inv_ext.y, (p.z - lo.z) * inv_ext.z};
    for (int ax = 0; ax < 3; ++ax) mlp_encode_axis(u[ax], out + ax * 2 * kMlpBands);
    mlp_encode_tail(w, n, a, out + 6 * kMlpBands);
}

}  // namespace xr
