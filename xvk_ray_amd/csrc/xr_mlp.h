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

// The 96 network inputs of a vertex at position p (in world units, scaled to
// [0,1]^3 by the scene box lo / inv_ext), outgoing direction w (unit,
// pointing away from the vertex), shading normal n, albedo a:
//   [0, 72)  sin(2^k pi u), cos(2^k pi u) for k = 0..11, u each of the three
//            normalised coordinates (axis-major: x's 24 values, then y, z)
//   [72, 75) w   [75, 78) n   [78, 81) a   [81] roughness (0: Lambertian)
//   [82, 96) zero
XR_HD void mlp_encode(v3 p, v3 w, v3 n, v3 a, v3 lo, v3 inv_ext, float* out) {
    const float u[3] = {(p.x - lo.x) * inv_ext.x, (p.y - lo.y) * inv_ext.y, (p.z - lo.z) * inv_ext.z};
    for (int ax = 0; ax < 3; ++ax) {
        float scale = XR_PI;
        for (int k = 0; k < kMlpBands; ++k) {
            float s, c;
            sincos_xr(scale * u[ax], &s, &c);
            out[ax * 2 * kMlpBands + 2 * k] = s;
            out[ax * 2 * kMlpBands + 2 * k + 1] = c;
            scale = scale * 2.0f;
        }
    }
    out[72] = w.x; out[73] = w.y; out[74] = w.z;
    out[75] = n.x; out[76] = n.y; out[77] = n.z;
    out[78] = a.x; out[79] = a.y; out[80] = a.z;
    out[81] = 0.0f;
    for (int i = 82; i < kMlpIn; ++i) out[i] = 0.0f;
}

// Per-frame selection of training pixels: pixel p trains when
// p % stride == (frame * 7919) % stride, so training slot k is pixel
// k * stride + offset.
XR_HD uint32_t mlp_train_offset(uint32_t frame, uint32_t stride) { return (frame * 7919u) % stride; }

}  // namespace xr
